/*
 * orc_solver.c -- oracle (test infrastructure only): preconditioners and
 * Krylov solvers for A(n) phi = b, standing in for the PETSc KSP that the
 * reference drives at potential.c:6-16 / hybrid.c:73-78.  FP64; see orc.h.
 */
#include "orc.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Host threads of the parallel loops (results do not depend on them). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

static int is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

/* In-place complex DFT of a strided line (radix-2 when n is a power of two,
 * direct summation otherwise); sign -1 forward, +1 inverse (unscaled). */
static void dft_line(double *re, double *im, int64_t n, double sign, double *tr, double *ti)
{
  if (is_pow2(n)) {
    for (int64_t i = 1, j = 0; i < n; i++) {
      int64_t bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) {
        double t = re[i]; re[i] = re[j]; re[j] = t;
        t = im[i]; im[i] = im[j]; im[j] = t;
      }
    }
    for (int64_t len = 2; len <= n; len <<= 1) {
      const double ang = sign * 2.0 * M_PI / (double)len;
      for (int64_t i = 0; i < n; i += len) {
        for (int64_t k = 0; k < len / 2; k++) {
          const double wr = cos(ang * (double)k), wi = sin(ang * (double)k);
          const int64_t a = i + k, b = i + k + len / 2;
          const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
          re[b] = re[a] - xr; im[b] = im[a] - xi;
          re[a] += xr; im[a] += xi;
        }
      }
    }
    return;
  }
  for (int64_t k = 0; k < n; k++) {
    double sr = 0.0, si = 0.0;
    for (int64_t j = 0; j < n; j++) {
      const double a = sign * 2.0 * M_PI * (double)((k * j) % n) / (double)n;
      sr += re[j] * cos(a) - im[j] * sin(a);
      si += re[j] * sin(a) + im[j] * cos(a);
    }
    tr[k] = sr; ti[k] = si;
  }
  memcpy(re, tr, sizeof(double) * n);
  memcpy(im, ti, sizeof(double) * n);
}

void orc_fft3d(const int64_t N[3], double *data, int inverse)
{
  const double sign = inverse ? 1.0 : -1.0;
  int64_t nmax = N[0];
  if (N[1] > nmax) nmax = N[1];
  if (N[2] > nmax) nmax = N[2];
  const int64_t str[3] = {1, N[0], N[0] * N[1]};
  const int64_t nc = N[0] * N[1] * N[2];
  for (int ax = 0; ax < 3; ax++) {
    const int64_t n = N[ax];
    /* the lines along ax are independent: one scratch set per thread */
#pragma omp parallel if (nc > 16384)
    {
      double *re = malloc(sizeof(double) * nmax), *im = malloc(sizeof(double) * nmax);
      double *tr = malloc(sizeof(double) * nmax), *ti = malloc(sizeof(double) * nmax);
#pragma omp for schedule(static)
      for (int64_t base = 0; base < nc; base++) {
        /* lines start where the index along ax is 0 */
        if ((base / str[ax]) % n != 0) continue;
        for (int64_t t = 0; t < n; t++) {
          re[t] = data[2 * (base + t * str[ax])];
          im[t] = data[2 * (base + t * str[ax]) + 1];
        }
        dft_line(re, im, n, sign, tr, ti);
        for (int64_t t = 0; t < n; t++) {
          data[2 * (base + t * str[ax])] = re[t];
          data[2 * (base + t * str[ax]) + 1] = im[t];
        }
      }
      free(re); free(im); free(tr); free(ti);
    }
  }
}

/* Row of the constant-coefficient operator A(nbar) (uniform density). */
static void uniform_row(int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                        double nbar, double v[19])
{
  (void)N;
  if (lhs_type == ORC_LHS_FULL) {
    /* evaluate the full row on a uniform 3x3x3 field (centre node 13) */
    const int64_t M[3] = {3, 3, 3};
    double n27[27];
    for (int q = 0; q < 27; q++) n27[q] = nbar;
    orc_lhs_full_row(M, d, K, n27, 13, v);
  } else if (lhs_type == ORC_LHS_LAPLACIAN) {
    const int64_t M[3] = {3, 3, 3};
    double n27[27] = {0};
    double A[27 * 27];
    orc_lhs_dense(ORC_LHS_LAPLACIAN, M, d, K, n27, A);
    int off[19][3];
    orc_stencil_offsets(off);
    for (int q = 0; q < 19; q++) {
      const int64_t col = (1 + off[q][0]) + 3 * ((1 + off[q][1]) + 3 * (1 + off[q][2]));
      v[q] = A[13 * 27 + col];
    }
  } else {
    for (int q = 0; q < 19; q++) v[q] = 0.0;
    v[18] = 1.0;
  }
}

/* Row c of the operator (any type). */
static void row_of(int lhs_type, const int64_t N[3], const double d[3], const double K[3], const double *n,
                   int64_t c, double v[19])
{
  if (lhs_type == ORC_LHS_FULL) orc_lhs_full_row(N, d, K, n, c, v);
  else uniform_row(lhs_type, N, d, K, 0.0, v);
}

/* Tridiagonal solve (Thomas algorithm): sub a, diagonal b, super c. */
static void thomas(int64_t m, const double *a, const double *b, const double *c, const double *r, double *x,
                   double *cp)
{
  double den = b[0];
  cp[0] = c[0] / den;
  x[0] = r[0] / den;
  for (int64_t k = 1; k < m; k++) {
    den = b[k] - a[k] * cp[k - 1];
    cp[k] = c[k] / den;
    x[k] = (r[k] - a[k] * x[k - 1]) / den;
  }
  for (int64_t k = m - 2; k >= 0; k--) x[k] -= cp[k] * x[k + 1];
}

/* Periodic (cyclic) tridiagonal solve by the Sherman-Morrison formula:
 * A = T + u v^T with u = (g, 0.., 0, c[m-1]), v = (1, 0.., 0, a[0]/g),
 * g = -b[0]; T differs from A in its two corner diagonal entries. */
static void cyclic_solve(int64_t m, const double *a, const double *b, const double *c, const double *r, double *x)
{
  if (m == 1) { x[0] = r[0] / b[0]; return; }
  if (m == 2) {  /* both neighbours are the other node */
    const double a00 = b[0], a01 = a[0] + c[0], a10 = a[1] + c[1], a11 = b[1];
    const double det = a00 * a11 - a01 * a10;
    x[0] = (r[0] * a11 - a01 * r[1]) / det;
    x[1] = (a00 * r[1] - a10 * r[0]) / det;
    return;
  }
  double *bb = malloc(sizeof(double) * (size_t)m * 4), *u = bb + m, *q = bb + 2 * m, *cp = bb + 3 * m;
  const double g = -b[0];
  for (int64_t k = 0; k < m; k++) { bb[k] = b[k]; u[k] = 0.0; }
  bb[0] = b[0] - g;
  bb[m - 1] = b[m - 1] - c[m - 1] * a[0] / g;
  u[0] = g;
  u[m - 1] = c[m - 1];
  thomas(m, a, bb, c, r, x, cp);
  thomas(m, a, bb, c, u, q, cp);
  const double f = (x[0] + a[0] / g * x[m - 1]) / (1.0 + q[0] + a[0] / g * q[m - 1]);
  for (int64_t k = 0; k < m; k++) x[k] -= f * q[k];
  free(bb);
}

static double dot(int64_t n, const double *a, const double *b);
static void residual(int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                     const double *n, const double *b, const double *phi, double *r);

/* x-y offset group (dx + 1) + 3 (dy + 1) of each of the 19 row entries */
static const int kGroup19[19] = {5, 3, 7, 1, 4, 4, 8, 2, 6, 0, 5, 5, 3, 3, 7, 7, 1, 1, 4};

/* z-means of the row entries grouped by x-y offset: C[o * nxy + col] */
static void zavg_coefficients(int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                              const double *n, double *C)
{
  const int64_t nxy = N[0] * N[1];
#pragma omp parallel for schedule(static) if (nxy > 1024)
  for (int64_t col = 0; col < nxy; col++) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t k = 0; k < N[2]; k++) {
      double v[19];
      row_of(lhs_type, N, d, K, n, col + nxy * k, v);
      for (int q = 0; q < 19; q++) acc[kGroup19[q]] += v[q];
    }
    for (int o = 0; o < 9; o++) C[o * nxy + col] = acc[o] / (double)N[2];
  }
}

static void a2_apply(const int64_t N[3], const double *C, const double *e, double *y)
{
  const int64_t nx = N[0], ny = N[1], nxy = nx * ny;
#pragma omp parallel for schedule(static) if (nxy > 16384)
  for (int64_t col = 0; col < nxy; col++) {
    const int64_t i = col % nx, j = col / nx;
    double acc = 0.0;
    for (int o = 0; o < 9; o++) {
      const int64_t ii = (i + (o % 3) - 1 + nx) % nx, jj = (j + (o / 3) - 1 + ny) % ny;
      acc += C[o * nxy + col] * e[ii + nx * jj];
    }
    y[col] = acc;
  }
}

/* 2-D FFT solve with the kz = 0 symbol of the mean-density operator */
static void f2_apply(const int64_t N[3], const double v[19], const double *r, double *z)
{
  const int64_t nx = N[0], ny = N[1], nxy = nx * ny;
  const int64_t N2[3] = {nx, ny, 1};
  int off[19][3];
  orc_stencil_offsets(off);
  double *buf = malloc(sizeof(double) * 2 * nxy);
  for (int64_t c = 0; c < nxy; c++) { buf[2 * c] = r[c]; buf[2 * c + 1] = 0.0; }
  orc_fft3d(N2, buf, 0);
  for (int64_t c = 0; c < nxy; c++) {
    const int64_t kx = c % nx, ky = c / nx;
    double lam = 0.0;
    for (int q = 0; q < 19; q++)
      lam += v[q] * cos(2.0 * M_PI * ((double)(kx * off[q][0]) / (double)nx + (double)(ky * off[q][1]) / (double)ny));
    if (c == 0 || lam == 0.0) { buf[2 * c] = buf[2 * c + 1] = 0.0; }
    else { buf[2 * c] /= lam; buf[2 * c + 1] /= lam; }
  }
  orc_fft3d(N2, buf, 1);
  for (int64_t c = 0; c < nxy; c++) z[c] = buf[2 * c] / (double)nxy;
  free(buf);
}

/* m steps of flexible GMRES from zero on A2 e = r2 (no restart) */
static void gmres2_fixed(int m, const int64_t N[3], const double *C, const double v19[19], const double *r2,
                         double *e)
{
  const int64_t nxy = N[0] * N[1];
  double *V = calloc((size_t)(m + 1) * nxy, sizeof(double)), *Z = calloc((size_t)m * nxy, sizeof(double));
  double *w = malloc(sizeof(double) * nxy);
  double *H = calloc((size_t)(m + 1) * m, sizeof(double)), *h = malloc(sizeof(double) * (m + 1));
  double *h2 = malloc(sizeof(double) * (m + 1)), *g = calloc((size_t)m + 1, sizeof(double));
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m), *y = malloc(sizeof(double) * m);
  for (int64_t c = 0; c < nxy; c++) e[c] = 0.0;
  double beta = sqrt(dot(nxy, r2, r2));
  int jm = 0;
  if (beta > 0.0) {
    for (int64_t c = 0; c < nxy; c++) V[c] = r2[c] / beta;
    g[0] = beta;
    for (int j = 0; j < m; j++) {
      f2_apply(N, v19, V + nxy * j, Z + nxy * j);
      a2_apply(N, C, Z + nxy * j, w);
      for (int q = 0; q <= j; q++) h[q] = dot(nxy, V + nxy * q, w);
      for (int q = 0; q <= j; q++) for (int64_t c = 0; c < nxy; c++) w[c] -= h[q] * V[nxy * q + c];
      for (int q = 0; q <= j; q++) h2[q] = dot(nxy, V + nxy * q, w);
      for (int q = 0; q <= j; q++) for (int64_t c = 0; c < nxy; c++) w[c] -= h2[q] * V[nxy * q + c];
      for (int q = 0; q <= j; q++) h[q] += h2[q];
      h[j + 1] = sqrt(dot(nxy, w, w));
      for (int64_t c = 0; c < nxy; c++) V[nxy * (j + 1) + c] = h[j + 1] > 0.0 ? w[c] / h[j + 1] : 0.0;
      for (int q = 0; q < j; q++) {
        const double t = cs[q] * h[q] + sn[q] * h[q + 1];
        h[q + 1] = -sn[q] * h[q] + cs[q] * h[q + 1];
        h[q] = t;
      }
      const double rr = sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
      cs[j] = rr > 0.0 ? h[j] / rr : 1.0;
      sn[j] = rr > 0.0 ? h[j + 1] / rr : 0.0;
      h[j] = rr;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int q = 0; q <= j; q++) H[q * m + j] = h[q];
      jm = j + 1;
      if (!(h[j + 1] > 0.0)) break;
    }
    for (int q = jm - 1; q >= 0; q--) {
      double s = g[q];
      for (int r = q + 1; r < jm; r++) s -= H[q * m + r] * y[r];
      y[q] = H[q * m + q] != 0.0 ? s / H[q * m + q] : 0.0;
    }
    for (int q = 0; q < jm; q++) for (int64_t c = 0; c < nxy; c++) e[c] += y[q] * Z[nxy * q + c];
  }
  free(V); free(Z); free(w); free(H); free(h); free(h2); free(g); free(cs); free(sn); free(y);
}

/* ---- xy-semicoarsening multigrid with a z-line smoother ---------------------
 * A V-cycle for the strongly z-coupled operator (parallel mobility about
 * 7.7e3 times the perpendicular one): the z-line block Jacobi smoother solves
 * every z-line exactly and damps the high x-y frequencies, and the coarse
 * grids halve x and y only (semicoarsening), keeping all z planes.  Level l
 * has N_l = (N_x / 2^l, N_y / 2^l, N_z) nodes and spacings (2^l dx, 2^l dy, dz);
 * its operator is the same 19-point operator rediscretised from the density
 * restricted to it.  Restriction is x-y full weighting (1/4, 1/2, 1/4 per
 * axis, periodic), prolongation its bilinear counterpart.  One V(2,2) cycle
 * from zero, damping 0.8, and 8 smoothing sweeps on the coarsest level. */
typedef struct {
  int nlev;
  int64_t N[ORC_MG_MAX_LEVELS][3];
  double d[ORC_MG_MAX_LEVELS][3];
  double *n[ORC_MG_MAX_LEVELS];
} mg_levels;

int orc_mg_levels(const int64_t N[3], int64_t out[ORC_MG_MAX_LEVELS][3])
{
  int64_t nx = N[0], ny = N[1];
  int L = 1;
  out[0][0] = nx; out[0][1] = ny; out[0][2] = N[2];
  while (L < ORC_MG_MAX_LEVELS && nx % 2 == 0 && ny % 2 == 0 && nx / 2 >= 3 && ny / 2 >= 3) {
    nx /= 2; ny /= 2;
    out[L][0] = nx; out[L][1] = ny; out[L][2] = N[2];
    L++;
  }
  return L;
}

/* c = x-y full weighting of f (f on Nf, c on Nc = (Nf_x / 2, Nf_y / 2, Nf_z)) */
void orc_mg_restrict(const int64_t Nf[3], const double *f, double *c)
{
  static const double w[3] = {0.25, 0.5, 0.25};
  const int64_t nx = Nf[0], ny = Nf[1], nz = Nf[2], cx = nx / 2, cy = ny / 2;
  for (int64_t k = 0; k < nz; k++)
    for (int64_t J = 0; J < cy; J++)
      for (int64_t I = 0; I < cx; I++) {
        double s = 0.0;
        for (int b = 0; b < 3; b++) {
          const int64_t j = (2 * J + b - 1 + ny) % ny;
          for (int a = 0; a < 3; a++) {
            const int64_t i = (2 * I + a - 1 + nx) % nx;
            s += (w[b] * w[a]) * f[i + nx * (j + ny * k)];
          }
        }
        c[I + cx * (J + cy * k)] = s;
      }
}

/* f += bilinear x-y interpolation of c (f on Nf) */
void orc_mg_prolong_add(const int64_t Nf[3], const double *c, double *f)
{
  const int64_t nx = Nf[0], ny = Nf[1], nz = Nf[2], cx = nx / 2, cy = ny / 2;
  for (int64_t k = 0; k < nz; k++)
    for (int64_t j = 0; j < ny; j++)
      for (int64_t i = 0; i < nx; i++) {
        const int64_t I0 = i / 2, J0 = j / 2;
        const int64_t I1 = (i % 2) ? (I0 + 1) % cx : I0, J1 = (j % 2) ? (J0 + 1) % cy : J0;
        const double *ck = c + cx * cy * k;
        const double v = 0.25 * ((ck[I0 + cx * J0] + ck[I1 + cx * J0]) + (ck[I0 + cx * J1] + ck[I1 + cx * J1]));
        f[i + nx * (j + ny * k)] += v;
      }
}

/* x += omega zline^-1 (b - A x), nsweep times */
static void mg_smooth(int lhs_type, const int64_t N[3], const double d[3], const double K[3], const double *n,
                      const double *b, double *x, int nsweep, double *r, double *e)
{
  const int64_t nc = N[0] * N[1] * N[2];
  for (int s = 0; s < nsweep; s++) {
    residual(lhs_type, N, d, K, n, b, x, r);
    orc_pc_apply(ORC_PC_ZLINE, lhs_type, N, d, K, n, r, e);
    for (int64_t c = 0; c < nc; c++) x[c] += ORC_MG_OMEGA * e[c];
  }
}

static void mg_vcycle(int lhs_type, const mg_levels *M, const double K[3], int l, const double *b, double *x)
{
  const int64_t *N = M->N[l];
  const int64_t nc = N[0] * N[1] * N[2];
  double *r = malloc(sizeof(double) * nc), *e = malloc(sizeof(double) * nc);
  for (int64_t c = 0; c < nc; c++) x[c] = 0.0;
  if (l == M->nlev - 1) {
    mg_smooth(lhs_type, N, M->d[l], K, M->n[l], b, x, ORC_MG_COARSE_SWEEPS, r, e);
  } else {
    mg_smooth(lhs_type, N, M->d[l], K, M->n[l], b, x, ORC_MG_SWEEPS, r, e);
    residual(lhs_type, N, M->d[l], K, M->n[l], b, x, r);
    const int64_t *Nc = M->N[l + 1];
    const int64_t ncc = Nc[0] * Nc[1] * Nc[2];
    double *bc = malloc(sizeof(double) * ncc), *xc = malloc(sizeof(double) * ncc);
    orc_mg_restrict(N, r, bc);
    mg_vcycle(lhs_type, M, K, l + 1, bc, xc);
    orc_mg_prolong_add(N, xc, x);
    free(bc); free(xc);
    mg_smooth(lhs_type, N, M->d[l], K, M->n[l], b, x, ORC_MG_SWEEPS, r, e);
  }
  free(r); free(e);
}

static void mg_apply(int lhs_type, const int64_t N[3], const double d[3], const double K[3], const double *n,
                     const double *r, double *z)
{
  mg_levels M;
  int64_t NL[ORC_MG_MAX_LEVELS][3];
  M.nlev = orc_mg_levels(N, NL);
  for (int l = 0; l < M.nlev; l++) {
    for (int a = 0; a < 3; a++) M.N[l][a] = NL[l][a];
    M.d[l][0] = d[0] * (double)(N[0] / NL[l][0]);
    M.d[l][1] = d[1] * (double)(N[1] / NL[l][1]);
    M.d[l][2] = d[2];
    M.n[l] = malloc(sizeof(double) * NL[l][0] * NL[l][1] * NL[l][2]);
    if (l == 0) memcpy(M.n[0], n, sizeof(double) * N[0] * N[1] * N[2]);
    else orc_mg_restrict(M.N[l - 1], M.n[l - 1], M.n[l]);
  }
  mg_vcycle(lhs_type, &M, K, 0, r, z);
  for (int l = 0; l < M.nlev; l++) free(M.n[l]);
  orc_remove_mean(N[0] * N[1] * N[2], z);
}

void orc_pc_apply(int pc, int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                  const double *n, const double *r, double *z)
{
  orc_pc_apply_deg(pc, lhs_type, N, d, K, n, r, z, ORC_CHEBY_DEGREE);
}

void orc_pc_apply_deg(int pc, int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                      const double *n, const double *r, double *z, int degree)
{
  const int64_t nc = N[0] * N[1] * N[2];
  if (pc == ORC_PC_MG) {
    mg_apply(lhs_type, N, d, K, n, r, z);
    return;
  }
  if (pc == ORC_PC_FFT2Z) {
    double *w = malloc(sizeof(double) * nc), *z2 = malloc(sizeof(double) * nc);
    orc_pc_apply_deg(ORC_PC_FFT2, lhs_type, N, d, K, n, r, z, degree);
    orc_lhs_apply(lhs_type, N, d, K, n, z, w);
    for (int64_t c = 0; c < nc; c++) w[c] = r[c] - w[c];
    orc_pc_apply_deg(ORC_PC_ZLINE, lhs_type, N, d, K, n, w, z2, degree);
    for (int64_t c = 0; c < nc; c++) z[c] += z2[c];
    free(w); free(z2);
    orc_remove_mean(nc, z);
    return;
  }
  if (pc == ORC_PC_FFT2) {
    const int64_t nxy = N[0] * N[1];
    orc_pc_apply_deg(ORC_PC_FFT, lhs_type, N, d, K, n, r, z, degree);
    double *w = malloc(sizeof(double) * nc), *C = malloc(sizeof(double) * 9 * nxy);
    double *r2 = calloc((size_t)nxy, sizeof(double)), *e2 = malloc(sizeof(double) * nxy);
    orc_lhs_apply(lhs_type, N, d, K, n, z, w);
#pragma omp parallel for schedule(static) if (nxy > 1024)
    for (int64_t col = 0; col < nxy; col++) {
      double s = 0.0;
      for (int64_t k = 0; k < N[2]; k++) s += r[col + nxy * k] - w[col + nxy * k];
      r2[col] = s / (double)N[2];
    }
    /* constants are not in the range of A2 (its columns sum to zero) */
    orc_remove_mean(nxy, r2);
    zavg_coefficients(lhs_type, N, d, K, n, C);
    double sn = 0.0;
    for (int64_t c = 0; c < nc; c++) sn += n[c];
    double v19[19];
    uniform_row(lhs_type, N, d, K, sn / (double)nc, v19);
    gmres2_fixed(ORC_FFT2_ITS, N, C, v19, r2, e2);
    for (int64_t c = 0; c < nc; c++) z[c] += e2[c % nxy];
    free(w); free(C); free(r2); free(e2);
    orc_remove_mean(nc, z);
    return;
  }
  if (pc == ORC_PC_ZLINE) {
    const int64_t m = N[2], plane = N[0] * N[1];
    /* independent columns: one scratch set per thread */
#pragma omp parallel if (nc > 16384)
    {
      double *buf = malloc(sizeof(double) * (size_t)m * 5);
      double *a = buf, *b = buf + m, *c = buf + 2 * m, *rr = buf + 3 * m, *x = buf + 4 * m;
#pragma omp for schedule(static)
      for (int64_t col = 0; col < plane; col++) {
        for (int64_t k = 0; k < m; k++) {
          double v[19];
          row_of(lhs_type, N, d, K, n, col + plane * k, v);
          a[k] = v[5];   /* (0, 0, -1) */
          b[k] = v[18];
          c[k] = v[4];   /* (0, 0, +1) */
          rr[k] = r[col + plane * k];
        }
        cyclic_solve(m, a, b, c, rr, x);
        for (int64_t k = 0; k < m; k++) z[col + plane * k] = x[k];
      }
      free(buf);
    }
    orc_remove_mean(nc, z);
    return;
  }
  if (pc == ORC_PC_CHEBY) {
    /* Chebyshev iteration (three-term recurrence) for D^-1 A on [lo, hi] */
    const double lo = 0.2, hi = 2.2, theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double *dg = malloc(sizeof(double) * (size_t)nc * 4), *res = dg + nc, *dv = dg + 2 * nc, *Ad = dg + 3 * nc;
    for (int64_t c = 0; c < nc; c++) {
      double v[19];
      row_of(lhs_type, N, d, K, n, c, v);
      dg[c] = v[18];
      res[c] = r[c] / dg[c];
      dv[c] = res[c] / theta;
      z[c] = dv[c];
    }
    double rho = 1.0 / sigma;
    for (int k = 1; k < degree; k++) {
      orc_lhs_apply(lhs_type, N, d, K, n, dv, Ad);
      const double rho1 = 1.0 / (2.0 * sigma - rho);
      const double c1 = rho1 * rho, c2 = 2.0 * rho1 / delta;
      for (int64_t c = 0; c < nc; c++) {
        res[c] -= Ad[c] / dg[c];
        dv[c] = c1 * dv[c] + c2 * res[c];
        z[c] += dv[c];
      }
      rho = rho1;
    }
    free(dg);
    orc_remove_mean(nc, z);
    return;
  }
  if (pc == ORC_PC_JACOBI) {
    for (int64_t c = 0; c < nc; c++) {
      double v[19];
      if (lhs_type == ORC_LHS_FULL) orc_lhs_full_row(N, d, K, n, c, v);
      else uniform_row(lhs_type, N, d, K, 0.0, v);
      z[c] = r[c] / v[18];
    }
    orc_remove_mean(nc, z);
    return;
  }
  if (pc == ORC_PC_FFT) {
    double s = 0.0;
    for (int64_t c = 0; c < nc; c++) s += n[c];
    const double nbar = s / (double)nc;
    double v[19];
    int off[19][3];
    uniform_row(lhs_type, N, d, K, nbar, v);
    orc_stencil_offsets(off);
    double *buf = malloc(sizeof(double) * 2 * nc);
    for (int64_t c = 0; c < nc; c++) { buf[2 * c] = r[c]; buf[2 * c + 1] = 0.0; }
    orc_fft3d(N, buf, 0);
#pragma omp parallel for schedule(static) if (nc > 16384)
    for (int64_t c = 0; c < nc; c++) {
      const int64_t kx = c % N[0], ky = (c / N[0]) % N[1], kz = c / (N[0] * N[1]);
      double lam = 0.0;
      for (int q = 0; q < 19; q++) {
        const double th = 2.0 * M_PI * ((double)(kx * off[q][0]) / (double)N[0] +
                                        (double)(ky * off[q][1]) / (double)N[1] +
                                        (double)(kz * off[q][2]) / (double)N[2]);
        lam += v[q] * cos(th);
      }
      if (c == 0 || lam == 0.0) { buf[2 * c] = buf[2 * c + 1] = 0.0; }
      else { buf[2 * c] /= lam; buf[2 * c + 1] /= lam; }
    }
    orc_fft3d(N, buf, 1);
    for (int64_t c = 0; c < nc; c++) z[c] = buf[2 * c] / (double)nc;
    free(buf);
    orc_remove_mean(nc, z);
    return;
  }
  memcpy(z, r, sizeof(double) * nc);
  orc_remove_mean(nc, z);
}

static double dot(int64_t n, const double *a, const double *b)
{
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static void residual(int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                     const double *n, const double *b, const double *phi, double *r)
{
  const int64_t nc = N[0] * N[1] * N[2];
  orc_lhs_apply(lhs_type, N, d, K, n, phi, r);
#pragma omp parallel for schedule(static) if (nc > 65536)
  for (int64_t c = 0; c < nc; c++) r[c] = b[c] - r[c];
}

static void gmres(int pc, int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                  const double *n, const double *b, double *phi, double tol,
                  int64_t max_it, int64_t m, int64_t info[2])
{
  const int64_t nc = N[0] * N[1] * N[2];
  double *V = malloc(sizeof(double) * nc * (m + 1));
  double *Z = malloc(sizeof(double) * nc * m);
  double *w = malloc(sizeof(double) * nc);
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
  double *g = malloc(sizeof(double) * (m + 1)), *y = malloc(sizeof(double) * m);
  double *h = malloc(sizeof(double) * (m + 1)), *h2 = malloc(sizeof(double) * (m + 1));
  int64_t it = 0;
  info[1] = 0;
  for (;;) {
    residual(lhs_type, N, d, K, n, b, phi, V);
    double beta = sqrt(dot(nc, V, V));
    if (beta <= tol) { info[1] = 1; break; }
    if (it >= max_it) break;
    for (int64_t c = 0; c < nc; c++) V[c] /= beta;
    for (int64_t q = 0; q <= m; q++) g[q] = 0.0;
    g[0] = beta;
    int64_t j = 0;
    for (; j < m && it < max_it; j++, it++) {
      double *zj = Z + nc * j;
      orc_pc_apply(pc, lhs_type, N, d, K, n, V + nc * j, zj);
      orc_lhs_apply(lhs_type, N, d, K, n, zj, w);
      /* two-pass classical Gram-Schmidt */
      /* the updates of each element in q order, elements in parallel */
      for (int64_t q = 0; q <= j; q++) h[q] = dot(nc, V + nc * q, w);
#pragma omp parallel for schedule(static) if (nc > 65536)
      for (int64_t c = 0; c < nc; c++) for (int64_t q = 0; q <= j; q++) w[c] -= h[q] * V[nc * q + c];
      for (int64_t q = 0; q <= j; q++) h2[q] = dot(nc, V + nc * q, w);
#pragma omp parallel for schedule(static) if (nc > 65536)
      for (int64_t c = 0; c < nc; c++) for (int64_t q = 0; q <= j; q++) w[c] -= h2[q] * V[nc * q + c];
      for (int64_t q = 0; q <= j; q++) h[q] += h2[q];
      h[j + 1] = sqrt(dot(nc, w, w));
      for (int64_t c = 0; c < nc; c++) V[nc * (j + 1) + c] = (h[j + 1] > 0.0) ? w[c] / h[j + 1] : 0.0;
      for (int64_t q = 0; q < j; q++) {
        const double t = cs[q] * h[q] + sn[q] * h[q + 1];
        h[q + 1] = -sn[q] * h[q] + cs[q] * h[q + 1];
        h[q] = t;
      }
      const double rr = sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
      cs[j] = (rr > 0.0) ? h[j] / rr : 1.0;
      sn[j] = (rr > 0.0) ? h[j + 1] / rr : 0.0;
      h[j] = rr;
      h[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int64_t q = 0; q <= j; q++) H[q * m + j] = h[q];
      if (fabs(g[j + 1]) <= tol) { j++; it++; break; }
    }
    for (int64_t q = j - 1; q >= 0; q--) {
      double s = g[q];
      for (int64_t r = q + 1; r < j; r++) s -= H[q * m + r] * y[r];
      y[q] = s / H[q * m + q];
    }
#pragma omp parallel for schedule(static) if (nc > 65536)
    for (int64_t c = 0; c < nc; c++) for (int64_t q = 0; q < j; q++) phi[c] += y[q] * Z[nc * q + c];
  }
  info[0] = it;
  free(V); free(Z); free(w); free(H); free(cs); free(sn); free(g); free(y); free(h); free(h2);
}

static void bicgstab(int pc, int lhs_type, const int64_t N[3], const double d[3], const double K[3],
                     const double *n, const double *b, double *phi, double tol,
                     int64_t max_it, int64_t info[2])
{
  const int64_t nc = N[0] * N[1] * N[2];
  double *r = malloc(sizeof(double) * nc), *rh = malloc(sizeof(double) * nc);
  double *p = calloc(nc, sizeof(double)), *v = calloc(nc, sizeof(double));
  double *yv = malloc(sizeof(double) * nc), *s = malloc(sizeof(double) * nc);
  double *zv = malloc(sizeof(double) * nc), *t = malloc(sizeof(double) * nc);
  residual(lhs_type, N, d, K, n, b, phi, r);
  memcpy(rh, r, sizeof(double) * nc);
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  int64_t it = 0;
  info[1] = 0;
  while (it < max_it) {
    if (sqrt(dot(nc, r, r)) <= tol) { info[1] = 1; break; }
    const double rho1 = dot(nc, rh, r);
    const double beta = (rho1 / rho) * (alpha / omega);
    for (int64_t c = 0; c < nc; c++) p[c] = r[c] + beta * (p[c] - omega * v[c]);
    orc_pc_apply(pc, lhs_type, N, d, K, n, p, yv);
    orc_lhs_apply(lhs_type, N, d, K, n, yv, v);
    alpha = rho1 / dot(nc, rh, v);
    for (int64_t c = 0; c < nc; c++) s[c] = r[c] - alpha * v[c];
    it++;
    if (sqrt(dot(nc, s, s)) <= tol) {
      for (int64_t c = 0; c < nc; c++) phi[c] += alpha * yv[c];
      info[1] = 1;
      break;
    }
    orc_pc_apply(pc, lhs_type, N, d, K, n, s, zv);
    orc_lhs_apply(lhs_type, N, d, K, n, zv, t);
    omega = dot(nc, t, s) / dot(nc, t, t);
    for (int64_t c = 0; c < nc; c++) {
      phi[c] += alpha * yv[c] + omega * zv[c];
      r[c] = s[c] - omega * t[c];
    }
    rho = rho1;
  }
  info[0] = it;
  free(r); free(rh); free(p); free(v); free(yv); free(s); free(zv); free(t);
}

void orc_solve(int ksp, int pc, int lhs_type, const int64_t N[3], const double d[3],
               const double K[3], const double *n, const double *b, double *phi,
               double rtol, double atol, int64_t max_it, int64_t restart,
               int64_t info[2], double resid[1])
{
  const int64_t nc = N[0] * N[1] * N[2];
  const double bn = sqrt(dot(nc, b, b));
  double tol = rtol * bn;
  if (atol > tol) tol = atol;
  if (ksp == ORC_KSP_BCGS) bicgstab(pc, lhs_type, N, d, K, n, b, phi, tol, max_it, info);
  else gmres(pc, lhs_type, N, d, K, n, b, phi, tol, max_it, restart, info);
  orc_remove_mean(nc, phi);
  double *r = malloc(sizeof(double) * nc);
  residual(lhs_type, N, d, K, n, b, phi, r);
  resid[0] = (bn > 0.0) ? sqrt(dot(nc, r, r)) / bn : sqrt(dot(nc, r, r));
  free(r);
}
