// chy_mg.hip -- x-y semicoarsening multigrid preconditioner with a z-line
// smoother (-pc_type mg; PETSc's -pc_type gamg / hypre / ml map here).
//
// The electron operator couples z about 7.7e3 times more strongly than x and
// y (B along z, kappa_e = 88).  The smoother solves every z-line block of
// A(n) exactly (the cyclic tridiagonal systems of the fft2z correction,
// factored once per solve) and, damped by 0.8, removes the high x-y
// frequencies of the error; the coarse grids halve x and y only and keep every
// z plane, so the error the smoother leaves (smooth in x and y) is what they
// represent.  Level l has (N_x / 2^l, N_y / 2^l, N_z) nodes and spacings
// (2^l dx, 2^l dy, dz); its operator is the 19-point operator rediscretised
// from the density restricted to it (x-y full weighting), so every level
// runs the same stencil and z-line kernels as the fine grid.  One V(2,2)
// cycle from a zero guess per application, 8 sweeps on the coarsest level.
// Unlike the FFT preconditioners it needs no power-of-two axes, and its
// iteration counts do not grow with the grid or with the density contrast.
#include "chy_internal.h"

namespace chy {

namespace {

constexpr int kMgMax = 8;
constexpr int kMgSweeps = 2;
constexpr int kMgCoarseSweeps = 8;
constexpr double kMgOmega = 0.8;

struct MgLevel {
  Grid g;
  double d[3];
  void *n = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *zf = nullptr;  // T arrays
  double *zcol = nullptr;  // per-column z-line scalars (2 per column)
};

struct Mg {
  int nlev = 0;
  MgLevel lv[kMgMax];
};

// c = x-y full weighting of f (weights 1/4, 1/2, 1/4 along x and y, periodic).
template <typename T>
__global__ __launch_bounds__(256) void k_mg_restrict(Grid gf, Grid gc, const T *__restrict__ f, T *__restrict__ c)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= gc.nc) return;
  const int I = (int)(q % gc.nx), J = (int)((q / gc.nx) % gc.ny), k = (int)(q / ((int64_t)gc.nx * gc.ny));
  const double w[3] = {0.25, 0.5, 0.25};
  double s = 0.0;
#pragma unroll
  for (int bq = 0; bq < 3; bq++) {
    const int j = wrapi(2 * J + bq - 1, gf.ny);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int i = wrapi(2 * I + a - 1, gf.nx);
      s += (w[bq] * w[a]) * (double)f[(int64_t)i + (int64_t)gf.nx * ((int64_t)j + (int64_t)gf.ny * k)];
    }
  }
  c[q] = (T)s;
}

// f += bilinear x-y interpolation of c.
template <typename T>
__global__ __launch_bounds__(256) void k_mg_prolong(Grid gf, Grid gc, const T *__restrict__ c, T *__restrict__ f)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= gf.nc) return;
  const int i = (int)(q % gf.nx), j = (int)((q / gf.nx) % gf.ny), k = (int)(q / ((int64_t)gf.nx * gf.ny));
  const int I0 = i >> 1, J0 = j >> 1;
  const int I1 = (i & 1) ? (I0 + 1 == gc.nx ? 0 : I0 + 1) : I0, J1 = (j & 1) ? (J0 + 1 == gc.ny ? 0 : J0 + 1) : J0;
  const T *ck = c + (int64_t)gc.nx * gc.ny * k;
  const double v = 0.25 * (((double)ck[I0 + (int64_t)gc.nx * J0] + (double)ck[I1 + (int64_t)gc.nx * J0]) +
                           ((double)ck[I0 + (int64_t)gc.nx * J1] + (double)ck[I1 + (int64_t)gc.nx * J1]));
  f[q] = (T)((double)f[q] + v);
}

inline unsigned nb_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Level sizes: halve x and y while both stay even and at least 3 nodes.
int level_sizes(int64_t nx, int64_t ny, int64_t out[kMgMax][2])
{
  int L = 1;
  out[0][0] = nx; out[0][1] = ny;
  while (L < kMgMax && nx % 2 == 0 && ny % 2 == 0 && nx / 2 >= 3 && ny / 2 >= 3) {
    nx /= 2; ny /= 2;
    out[L][0] = nx; out[L][1] = ny;
    L++;
  }
  return L;
}

template <typename T>
void vcycle(chy_ctx *ctx, Mg &M, int l, const T *b, T *x)
{
  MgLevel &L = M.lv[l];
  const Grid &g = L.g;
  const int lt = ctx->prm.lhs_type;
  const T *n = (const T *)L.n;
  T *r = (T *)L.r, *zf = (T *)L.zf;
  // first sweep from x = 0: x = omega zline^-1 b (the thread-per-column solve
  // overwrites its input, so it gets a copy)
  launch_fill<T>(ctx, g.nc, x, (T)0);
  (void)hipMemcpyAsync(r, b, sizeof(T) * g.nc, hipMemcpyDeviceToDevice, ctx->stream);
  launch_zline_apply_coarse<T>(ctx, g, zf, L.zcol, r, x, kMgOmega);
  auto sweep = [&]() {
    launch_residual_g<T>(ctx, g, L.d, lt, n, b, x, r);
    launch_zline_apply_coarse<T>(ctx, g, zf, L.zcol, r, x, kMgOmega);
  };
  const bool coarsest = l == M.nlev - 1;
  for (int s = 1; s < (coarsest ? kMgCoarseSweeps : kMgSweeps); s++) sweep();
  if (coarsest) return;
  MgLevel &C = M.lv[l + 1];
  launch_residual_g<T>(ctx, g, L.d, lt, n, b, x, r);
  hipLaunchKernelGGL((k_mg_restrict<T>), nb_of(C.g.nc), kBlock, 0, ctx->stream, g, C.g, (const T *)r, (T *)C.b);
  vcycle<T>(ctx, M, l + 1, (const T *)C.b, (T *)C.x);
  hipLaunchKernelGGL((k_mg_prolong<T>), nb_of(g.nc), kBlock, 0, ctx->stream, g, C.g, (const T *)C.x, x);
  for (int s = 0; s < kMgSweeps; s++) sweep();
}

}  // namespace

int mg_setup(chy_ctx *ctx)
{
  const chy_params &p = ctx->prm;
  Mg *M = new Mg();
  ctx->mg = M;
  int64_t sz[kMgMax][2];
  M->nlev = level_sizes(p.N[0], p.N[1], sz);
  const size_t ts = ctx->mixed ? 4 : 8;
  for (int l = 0; l < M->nlev; l++) {
    MgLevel &L = M->lv[l];
    Grid &g = L.g;
    g = ctx->g;
    g.nx = (int)sz[l][0];
    g.ny = (int)sz[l][1];
    g.mx = fastdiv_magic((uint32_t)g.nx);
    g.my = fastdiv_magic((uint32_t)g.ny);
    g.nc = (int64_t)g.nx * g.ny * g.nz;
    L.d[0] = p.d[0] * (double)(p.N[0] / g.nx);
    L.d[1] = p.d[1] * (double)(p.N[1] / g.ny);
    L.d[2] = p.d[2];
    for (int a = 0; a < 3; a++) g.d[a] = L.d[a];
    const int64_t nc = g.nc, plane = (int64_t)g.nx * g.ny;
    unsigned char *q = nullptr;
    TRY(dalloc(ctx, &q, (int64_t)ts * nc)); L.r = q;
    TRY(dalloc(ctx, &q, (int64_t)ts * 4 * nc)); L.zf = q;
    TRY(dalloc(ctx, &L.zcol, 2 * plane));
    if (l > 0) {
      TRY(dalloc(ctx, &q, (int64_t)ts * nc)); L.n = q;
      TRY(dalloc(ctx, &q, (int64_t)ts * nc)); L.x = q;
      TRY(dalloc(ctx, &q, (int64_t)ts * nc)); L.b = q;
    }
  }
  return CHY_OK;
}

void mg_release(chy_ctx *ctx)
{
  Mg *M = (Mg *)ctx->mg;
  if (!M) return;
  for (int l = 0; l < M->nlev; l++) {
    MgLevel &L = M->lv[l];
    void *ps[] = {L.r, L.zf, L.zcol, l > 0 ? L.n : nullptr, L.x, L.b};
    for (void *p : ps) if (p) (void)hipFree(p);
  }
  delete M;
  ctx->mg = nullptr;
}

int mg_levels(const chy_ctx *ctx) { return ctx->mg ? ((const Mg *)ctx->mg)->nlev : 0; }

// Once per solve: the density of every level and the z-line factors of every
// level's operator.
template <typename T>
int mg_prepare(chy_ctx *ctx, const T *n)
{
  Mg *M = (Mg *)ctx->mg;
  if (!M) return set_err(ctx, CHY_ERR_LIB, "multigrid levels were not set up");
  M->lv[0].n = (void *)n;
  for (int l = 0; l < M->nlev; l++) {
    MgLevel &L = M->lv[l];
    if (l > 0)
      hipLaunchKernelGGL((k_mg_restrict<T>), nb_of(L.g.nc), kBlock, 0, ctx->stream, M->lv[l - 1].g, L.g,
                         (const T *)M->lv[l - 1].n, (T *)L.n);
    launch_zline_factor_g<T>(ctx, L.g, L.d, ctx->prm.lhs_type, (const T *)L.n, (T *)L.zf, L.zcol);
  }
  CHY_HIP(hipGetLastError());
  return CHY_OK;
}

// z = one V-cycle applied to r (level 0: the context's grid and density).
template <typename T>
int mg_apply(chy_ctx *ctx, const T *r, T *z)
{
  Mg *M = (Mg *)ctx->mg;
  if (!M || !M->lv[0].n) return set_err(ctx, CHY_ERR_LIB, "multigrid used before its set-up");
  vcycle<T>(ctx, *M, 0, r, z);
  CHY_HIP(hipGetLastError());
  return CHY_OK;
}

template int mg_prepare<double>(chy_ctx *, const double *);
template int mg_prepare<float>(chy_ctx *, const float *);
template int mg_apply<double>(chy_ctx *, const double *, double *);
template int mg_apply<float>(chy_ctx *, const float *, float *);

}  // namespace chy
