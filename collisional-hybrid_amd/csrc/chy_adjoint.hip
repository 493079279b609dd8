// chy_adjoint.hip -- reverse-mode derivative of the deterministic step map
// (FP64).  The forward map is the loop body of the reference driver
// (hybrid.c:101-123): deposit and potential solve of the initial ions, then
// per step gather + Boris + collisions + move (particles.c:347-718), stable
// sort, CIC deposit (particles.c:136-255), forcing (rhs.c:133-303) and the
// solve A(n) phi = b (lhs.c:350-543, potential.c:6-16).  The objective is the
// final wave energy J = sum_c (n_c - mean n)^2.
//
// Transposes used by the backward sweep:
//  deposit <-> gather   the same CIC weights; the adjoint of the deposit is a
//                       per-ion gather of the grid adjoints (plus the weight
//                       derivatives), the adjoint of the gather is a
//                       deterministic deposit of the per-ion field adjoints;
//  solve                A(n)^T mu = P lambda_phi with the same Krylov solver
//                       and preconditioner (A(n) and A(n)^T share the
//                       constant null space), then lambda_n -= d(mu^T A phi)/dn;
//  push                 Boris transposed step by step; the scattering of a
//                       colliding ion is differentiated in forward mode with
//                       its neutral velocity and angles frozen, and the
//                       collision decisions are those of the forward run
//                       (common random numbers);
//  move, sort           identity and a permutation.
// The forward run keeps a checkpoint (ions, segments, phi, collision
// normaliser, step counter) every `ckpt` steps; the backward sweep recomputes
// each segment from its checkpoint and walks it in reverse.
#include <cmath>
#include <vector>

#include "chy_internal.h"
#include "chy_push.h"

namespace chy {

namespace {

// Per-ion adjoint variables.
struct AdjP { double *x, *y, *z, *vx, *vy, *vz, *w; };

// ---- forward-mode numbers with three tangents (scattering Jacobian) ----------------
struct D3 { double v, d[3]; };
__device__ __forceinline__ D3 dconst(double a) { return D3{a, {0.0, 0.0, 0.0}}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}}; }
__device__ __forceinline__ D3 operator*(D3 a, D3 b)
{
  return D3{a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.v, {s * a.d[0], s * a.d[1], s * a.d[2]}}; }
__device__ __forceinline__ D3 operator/(D3 a, D3 b)
{
  const double q = a.v / b.v, ib = 1.0 / b.v;
  return D3{q, {(a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib, (a.d[2] - q * b.d[2]) * ib}};
}
__device__ __forceinline__ D3 dsqrt(D3 a)
{
  const double r = sqrt(a.v), h = r > 0.0 ? 0.5 / r : 0.0;
  return D3{r, {a.d[0] * h, a.d[1] * h, a.d[2] * h}};
}

// lambda <- J^T lambda for the scattering map v -> v' of an accepted
// collision (the kinematics of scatter() with the neutral and angles fixed).
__device__ void scatter_vjp(const PushConst &pc, uint32_t pid, double vx, double vy, double vz, double &lx,
                            double &ly, double &lz)
{
  double vnx, vny, vnz, ud;
  neutral_draw<double>(pc, pid, vnx, vny, vnz, ud);
  const U4 b = philox4x32_10(U4{pid, pc.step, STREAM_COLLIDE_B, 0u}, pc.k0, pc.k1);
  const double uz = 2.0 * u24<double>(b.x) - 1.0;
  const double uperp = sqrt(1.0 - uz * uz);
  const double uphi = 2.0 * M_PI * u24<double>(b.y);
  const double ux = uperp * cos(uphi), uy = uperp * sin(uphi);
  const D3 Vx{vx, {1.0, 0.0, 0.0}}, Vy{vy, {0.0, 1.0, 0.0}}, Vz{vz, {0.0, 0.0, 1.0}};
  const double mi = pc.mi, mn = pc.mn, M = mn + mi;
  const D3 vrx = Vx - dconst(vnx), vry = Vy - dconst(vny), vrz = Vz - dconst(vnz);
  const D3 vrr = dsqrt(vrx * vrx + vry * vry + vrz * vrz);
  const D3 vcx = (1.0 / M) * (mi * Vx + dconst(vnx * mn));
  const D3 vcy = (1.0 / M) * (mi * Vy + dconst(vny * mn));
  const D3 vcz = (1.0 / M) * (mi * Vz + dconst(vnz * mn));
  const D3 ex = Vx - vcx, ey = Vy - vcy, ez = Vz - vcz;
  const D3 vcr = dsqrt(ex * ex + ey * ey + ez * ez);
  const D3 costht = vrz / vrr;
  const D3 sintht = dsqrt(dconst(1.0) - costht * costht);
  D3 cosphi = dconst(1.0), sinphi = dconst(0.0);
  if (sintht.v > 1e-12) {
    cosphi = vrx / (vrr * sintht);
    sinphi = vry / (vrr * sintht);
  }
  const D3 ux1 = uz * sintht + ux * costht, uz1 = uz * costht - ux * sintht;
  const D3 ux2 = ux1 * cosphi - uy * sinphi, uy2 = ux1 * sinphi + uy * cosphi;
  const D3 fx = vcx + vcr * ux2, fy = vcy + vcr * uy2, fz = vcz + vcr * uz1;
  const double ax = lx, ay = ly, az = lz;
  lx = fx.d[0] * ax + fy.d[0] * ay + fz.d[0] * az;
  ly = fx.d[1] * ax + fy.d[1] * ay + fz.d[1] * az;
  lz = fx.d[2] * ax + fy.d[2] * ay + fz.d[2] * az;
}

// CIC weights of an ion and their derivatives along x, y, z (corner order
// q = dx + 2 dy + 4 dz, as the deposit and gather).
__device__ __forceinline__ void weights_and_slopes(const Grid &g, double x, double y, double z, int &i, int &j, int &k,
                                                   double W[8], double Wx[8], double Wy[8], double Wz[8])
{
  double fx, fy, fz;
  cell_coord(x, g.p0[0], g.d[0], g.nx, i, fx);
  cell_coord(y, g.p0[1], g.d[1], g.ny, j, fy);
  cell_coord(z, g.p0[2], g.d[2], g.nz, k, fz);
  const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, wz[2] = {1.0 - fz, fz};
  const double sx[2] = {-1.0 / g.d[0], 1.0 / g.d[0]}, sy[2] = {-1.0 / g.d[1], 1.0 / g.d[1]};
  const double sz[2] = {-1.0 / g.d[2], 1.0 / g.d[2]};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
    W[q] = (wz[c] * wy[b]) * wx[a];
    Wx[q] = (wz[c] * wy[b]) * sx[a];
    Wy[q] = (wz[c] * sy[b]) * wx[a];
    Wz[q] = (sz[c] * wy[b]) * wx[a];
  }
}

// Adjoint of the deposit n = sum w W, G = sum w W v for the ions of `st`:
// lambda_w += W.(ln + v.lG), lambda_v += w W.lG, lambda_x += w dW.(ln + v.lG).
__global__ __launch_bounds__(256) void k_adj_deposit(Grid g, int64_t np, PState<double> st, bool weighted,
                                                     const double *__restrict__ ln, const double *__restrict__ lgx,
                                                     const double *__restrict__ lgy, const double *__restrict__ lgz,
                                                     AdjP lam)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int i, j, k;
  double W[8], Wx[8], Wy[8], Wz[8];
  weights_and_slopes(g, st.x[p], st.y[p], st.z[p], i, j, k, W, Wx, Wy, Wz);
  const double vx = st.vx[p], vy = st.vy[p], vz = st.vz[p], w = weighted ? st.w[p] : 1.0;
  double gw = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int64_t c = lin(g, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2));
    const double a0 = lgx[c], a1 = lgy[c], a2 = lgz[c];
    const double h = ln[c] + (vx * a0 + vy * a1 + vz * a2);
    gw += W[q] * h;
    g0 += W[q] * a0; g1 += W[q] * a1; g2 += W[q] * a2;
    hx += Wx[q] * h; hy += Wy[q] * h; hz += Wz[q] * h;
  }
  lam.w[p] += gw;
  lam.vx[p] += w * g0; lam.vy[p] += w * g1; lam.vz[p] += w * g2;
  lam.x[p] += w * hx; lam.y[p] += w * hy; lam.z[p] += w * hz;
}

// Undo the sort (slot s came from pre-sort index perm[s]) and the move
// x' = x + v dt: lambda_x passes through, lambda_v gains dt lambda_x.
__global__ __launch_bounds__(256) void k_adj_unsort_move(int64_t np, const uint32_t *__restrict__ perm, double dt,
                                                         AdjP in, AdjP out)
{
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= np) return;
  const uint32_t i = perm[s];
  const double lx = in.x[s], ly = in.y[s], lz = in.z[s];
  out.x[i] = lx; out.y[i] = ly; out.z[i] = lz;
  out.vx[i] = in.vx[s] + dt * lx;
  out.vy[i] = in.vy[s] + dt * ly;
  out.vz[i] = in.vz[s] + dt * lz;
  out.w[i] = in.w[s];
}

// Adjoint of gather + Boris + collision for the ions of `st` (the state at
// the start of the step, in its own order).  lam holds lambda_x and the
// adjoint of the post-collision velocity on entry, lambda_x and lambda_v of
// the state on exit; lE receives the adjoint of each ion's gathered field.
__global__ __launch_bounds__(256) void k_adj_push(Grid g, PushConst pc, int64_t np, PState<double> st,
                                                  const Vec4<double> *__restrict__ E4, const double *__restrict__ Sdev,
                                                  bool collide, AdjP lam, double *__restrict__ lex,
                                                  double *__restrict__ ley, double *__restrict__ lez)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int i, j, k;
  double W[8], Wx[8], Wy[8], Wz[8];
  weights_and_slopes(g, st.x[p], st.y[p], st.z[p], i, j, k, W, Wx, Wy, Wz);
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  Vec4<double> En[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    En[q] = E4[lin(g, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2))];
    e0 += W[q] * En[q].x;
    e1 += W[q] * En[q].y;
    e2 += W[q] * En[q].z;
  }
  double lx = lam.vx[p], ly = lam.vy[p], lz = lam.vz[p];
  if (collide) {
    double vx = st.vx[p], vy = st.vy[p], vz = st.vz[p];
    boris<double>(pc, e0, e1, e2, vx, vy, vz);
    int capped = 0;
    if (collide_decide<double>(pc, Sdev[0], st.pid[p], vx, vy, vz, capped)) {
      double tx = vx, ty = vy, tz = vz;
      if (scatter<double>(pc, st.pid[p], tx, ty, tz) == 1) scatter_vjp(pc, st.pid[p], vx, vy, vz, lx, ly, lz);
    }
  }
  // Boris transposed: v* = vq + tau E, vq = vm + vp x s, vp = vm + vm x t, vm = v + tau E
  const double tau = pc.tau, t0 = pc.t[0], t1 = pc.t[1], t2 = pc.t[2], s0 = pc.s[0], s1 = pc.s[1], s2 = pc.s[2];
  double lE0 = tau * lx, lE1 = tau * ly, lE2 = tau * lz;
  const double lp0 = s1 * lz - s2 * ly, lp1 = s2 * lx - s0 * lz, lp2 = s0 * ly - s1 * lx;  // s x lambda_q
  const double lm0 = lx + lp0 + (t1 * lp2 - t2 * lp1);
  const double lm1 = ly + lp1 + (t2 * lp0 - t0 * lp2);
  const double lm2 = lz + lp2 + (t0 * lp1 - t1 * lp0);
  lE0 += tau * lm0; lE1 += tau * lm1; lE2 += tau * lm2;
  lam.vx[p] = lm0; lam.vy[p] = lm1; lam.vz[p] = lm2;
  // gather transposed for the position
  double hx = 0.0, hy = 0.0, hz = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const double h = En[q].x * lE0 + En[q].y * lE1 + En[q].z * lE2;
    hx += Wx[q] * h; hy += Wy[q] * h; hz += Wz[q] * h;
  }
  lam.x[p] += hx; lam.y[p] += hy; lam.z[p] += hz;
  lex[p] = lE0; ley[p] = lE1; lez[p] = lE2;
}

// Per-cell sums of a per-ion value over the cell's ions in sorted order.
__global__ __launch_bounds__(256) void k_cell_sum(int64_t nc, const uint32_t *__restrict__ offs,
                                                  const double *__restrict__ v, double *__restrict__ out)
{
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double s = 0.0;
  for (uint32_t q = offs[c]; q < offs[c + 1]; q++) s += v[q];
  out[c] = s;
}

// ln = 2 (n - mean n): gradient of the wave energy.
__global__ __launch_bounds__(256) void k_wave_grad(int64_t nc, const double *__restrict__ n,
                                                   const double *__restrict__ nsum, double *__restrict__ ln)
{
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  ln[c] = 2.0 * (n[c] - nsum[0] / (double)nc);
}

inline unsigned nb_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// One saved state of the forward run.
struct Stage {
  PState<double> st;
  uint32_t *offs = nullptr;  // cell segments of st (nc + 1)
  double *phi = nullptr;     // potential the next push uses
  double *S = nullptr;       // collision normaliser of st (device scalar)
  int64_t step = 0;          // collision stream counter
  bool ordered = false;      // st in the stable order of offs (the sort's neighbour-move path)
};

struct Tape {
  int64_t nsteps = 0, ckpt = 1, np = 0;
  bool done_forward = false;
  std::vector<Stage> ck;   // checkpoints at steps 0, ckpt, 2 ckpt, ...
  std::vector<Stage> seg;  // recomputed states 1..ckpt of one segment
  std::vector<uint32_t *> perm;  // final sort slot -> source index, per step of one segment
  Stage fin;               // state after the last step
  AdjP lam[2] = {};
  double *lE[3] = {nullptr, nullptr, nullptr};
  double *ln = nullptr, *lg = nullptr, *lphi = nullptr, *g3 = nullptr, *lEn = nullptr, *wrk = nullptr;
  std::vector<void *> extra;  // other allocations
};

int alloc_stage(chy_ctx *ctx, Stage &s, int64_t np)
{
  TRY(alloc_state<double>(ctx, s.st, np, ctx->weighted));
  TRY(dalloc(ctx, &s.offs, ctx->g.nc + 1));
  TRY(dalloc(ctx, &s.phi, ctx->g.nc));
  TRY(dalloc(ctx, &s.S, 1));
  return CHY_OK;
}

void free_stage(Stage &s)
{
  free_state(s.st);
  void *ps[] = {s.offs, s.phi, s.S};
  for (void *p : ps) if (p) (void)hipFree(p);
  s = Stage();
}

int copy_state(chy_ctx *ctx, const PState<double> &a, PState<double> &b, int64_t np)
{
  const size_t d = sizeof(double) * (size_t)np, u = sizeof(uint32_t) * (size_t)np;
  hipStream_t s = ctx->stream;
  CHY_HIP(hipMemcpyAsync(b.x, a.x, d, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.y, a.y, d, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.z, a.z, d, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.vx, a.vx, d, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.vy, a.vy, d, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.vz, a.vz, d, hipMemcpyDeviceToDevice, s));
  if (a.w && b.w) CHY_HIP(hipMemcpyAsync(b.w, a.w, d, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.pid, a.pid, u, hipMemcpyDeviceToDevice, s));
  CHY_HIP(hipMemcpyAsync(b.cell, a.cell, u, hipMemcpyDeviceToDevice, s));
  return CHY_OK;
}

// context state -> stage
int save_stage(chy_ctx *ctx, Stage &s)
{
  const int64_t nc = ctx->g.nc;
  TRY(copy_state(ctx, ctx->pd[ctx->cur], s.st, ctx->np));
  CHY_HIP(hipMemcpyAsync(s.offs, ctx->offsets, sizeof(uint32_t) * (nc + 1), hipMemcpyDeviceToDevice, ctx->stream));
  CHY_HIP(hipMemcpyAsync(s.phi, ctx->phi, sizeof(double) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  CHY_HIP(hipMemcpyAsync(s.S, ctx->S_dev, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  s.step = ctx->step;
  s.ordered = ctx->order_valid;
  return CHY_OK;
}

// stage -> context state
int restore_stage(chy_ctx *ctx, const Stage &s)
{
  const int64_t nc = ctx->g.nc;
  TRY(copy_state(ctx, s.st, ctx->pd[ctx->cur], ctx->np));
  CHY_HIP(hipMemcpyAsync(ctx->offsets, s.offs, sizeof(uint32_t) * (nc + 1), hipMemcpyDeviceToDevice, ctx->stream));
  CHY_HIP(hipMemcpyAsync(ctx->phi, s.phi, sizeof(double) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  CHY_HIP(hipMemcpyAsync(ctx->S_dev, s.S, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  ctx->step = s.step;
  ctx->order_valid = s.ordered;
  ctx->dir_valid = false;
  ctx->keys_moved = true;  // not known: the next push writes no directions
  ctx->partials_fresh = false;
  ctx->phi_prev_valid = false;
  ctx->phi_valid = true;
  ctx->moments_valid = false;
  return CHY_OK;
}

void free_tape(Tape *t)
{
  if (!t) return;
  for (Stage &s : t->ck) free_stage(s);
  for (Stage &s : t->seg) free_stage(s);
  free_stage(t->fin);
  for (uint32_t *p : t->perm) if (p) (void)hipFree(p);
  for (void *p : t->extra) if (p) (void)hipFree(p);
  delete t;
}

int alloc_adp(chy_ctx *ctx, Tape *t, AdjP &a, int64_t np)
{
  double **f[7] = {&a.x, &a.y, &a.z, &a.vx, &a.vy, &a.vz, &a.w};
  for (double **p : f) {
    TRY(dalloc(ctx, p, np));
    t->extra.push_back(*p);
  }
  return CHY_OK;
}

int grid_alloc(chy_ctx *ctx, Tape *t, double **p, int64_t n)
{
  TRY(dalloc(ctx, p, n));
  t->extra.push_back(*p);
  return CHY_OK;
}

int fetch_n(chy_ctx *ctx, const double *dev, int n, double *host)
{
  CHY_HIP(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  CHY_HIP(hipStreamSynchronize(ctx->stream));
  return CHY_OK;
}

double *mom64(chy_ctx *ctx, int a) { return (double *)ctx->mom[a]; }

// ln = 2 (n - mean n) of the current moments; returns J = sum (n - mean n)^2.
int wave_energy(chy_ctx *ctx, Tape *t, double *J)
{
  const int64_t nc = ctx->g.nc;
  launch_sum<double>(ctx, mom64(ctx, 0), nc, ctx->red_out);
  hipLaunchKernelGGL(k_wave_grad, nb_for(nc), kBlock, 0, ctx->stream, nc, mom64(ctx, 0), ctx->red_out, t->ln);
  launch_mdot<double>(ctx, 1, t->ln, nc, t->ln, nc, ctx->red_out + 1);
  double s;
  TRY(fetch_n(ctx, ctx->red_out + 1, 1, &s));
  *J = 0.25 * s;
  return CHY_OK;
}

// Backward through "moments of stage A, then phi = solve(A's moments)":
// lambda of A's ions += deposit^T (lambda_n, lambda_G), where lambda_n and
// lambda_G collect dJ/dn (final stage only) and the adjoint of the solve
// for lambda_phi (null for the final stage, whose phi J does not use).
int moments_adjoint(chy_ctx *ctx, Tape *t, const Stage &A, bool final, const double *lphi, AdjP lam, double lE0[3])
{
  const int64_t nc = ctx->g.nc;
  double *const mo[4] = {mom64(ctx, 0), mom64(ctx, 1), mom64(ctx, 2), mom64(ctx, 3)};
  deposit_state(ctx, A.st, A.offs, ctx->weighted, mo);
  ctx->moments_valid = true;
  double *lg0 = t->lg, *lg1 = t->lg + nc, *lg2 = t->lg + 2 * nc;
  if (final) {
    double J;
    TRY(wave_energy(ctx, t, &J));
  } else {
    launch_fill<double>(ctx, nc, t->ln, 0.0);
  }
  launch_fill<double>(ctx, 3 * nc, t->lg, 0.0);
  if (lphi) {
    // mu = A(n)^-T P lambda_phi (phi was projected onto mean zero)
    CHY_HIP(hipMemcpyAsync(ctx->b, lphi, sizeof(double) * nc, hipMemcpyDeviceToDevice, ctx->stream));
    launch_sum<double>(ctx, ctx->b, nc, ctx->red_out);
    launch_shift<double>(ctx, nc, ctx->b, ctx->red_out, -1.0 / (double)nc);
    launch_fill<double>(ctx, nc, ctx->phi, 0.0);
    ctx->phi_valid = true;
    ctx->solve_transpose = true;
    chy_solve_stats st;
    const int rc = ctx_solve(ctx, &st);
    ctx->solve_transpose = false;
    if (rc) return rc;
    // an unconverged mu would enter every later adjoint silently
    if (!st.converged) {
      char m[160];
      snprintf(m, sizeof m, "adjoint: the transposed potential solve did not converge (relative residual %.3e after "
               "%lld iterations)", st.residual, (long long)st.iterations);
      return set_err(ctx, CHY_ERR_NOT_CONVERGED, m);
    }
    const double *mu = ctx->phi;
    // phi depends on n through A(n): lambda_n -= d(mu^T A(n) phi_A)/dn
    launch_lhs_dn(ctx, mu, A.phi, -1.0, t->ln);
    // and through b(n, G, E0)
    launch_rhs_full_adj(ctx, mu, t->ln, lg0, lg1, lg2);
    launch_grad3(ctx, mo[0], t->g3);
    launch_mdot<double>(ctx, 3, t->g3, nc, mu, nc, ctx->red_out);
    double dn[3];
    TRY(fetch_n(ctx, ctx->red_out, 3, dn));
    const RhsCoef rc_ = rhs_coef(ctx);
    for (int b = 0; b < 3; b++)
      lE0[b] += rc_.scale * (rc_.r[0][b] * dn[0] + rc_.r[1][b] * dn[1] + rc_.r[2][b] * dn[2]);
  }
  hipLaunchKernelGGL(k_adj_deposit, nb_for(ctx->np), kBlock, 0, ctx->stream, ctx->g, ctx->np, A.st, ctx->weighted,
                     t->ln, lg0, lg1, lg2, lam);
  return CHY_OK;
}

// Backward through the push of step k: ions of stage B (start of the step),
// `perm` the sort of that step, lam_in the adjoint of the sorted result,
// lam_out that of B.  Leaves lambda_phi of B's potential in t->lphi.
int push_adjoint(chy_ctx *ctx, Tape *t, const Stage &B, const uint32_t *perm, AdjP lam_in, AdjP lam_out,
                 double lE0[3])
{
  const int64_t nc = ctx->g.nc, np = ctx->np;
  const chy_params &p = ctx->prm;
  hipLaunchKernelGGL(k_adj_unsort_move, nb_for(np), kBlock, 0, ctx->stream, np, perm, p.dt, lam_in, lam_out);
  launch_efield<double>(ctx, B.phi, (double *)ctx->efield[0]);
  ctx->step = B.step;
  PushConst pc = push_const(ctx);
  const bool collide = p.mn > 0.0;
  if (collide) pc.Nc = (double)collision_count(p, np);
  hipLaunchKernelGGL(k_adj_push, nb_for(np), kBlock, 0, ctx->stream, ctx->g, pc, np, B.st,
                     (const Vec4<double> *)ctx->efield[0], B.S, collide, lam_out, t->lE[0], t->lE[1], t->lE[2]);
  // lambda_E at the nodes: deterministic deposit of the per-ion lambda_E
  PState<double> q = B.st;
  q.vx = t->lE[0]; q.vy = t->lE[1]; q.vz = t->lE[2]; q.w = nullptr;
  double *const out[4] = {t->lEn, t->lEn + nc, t->lEn + 2 * nc, t->lEn + 3 * nc};
  deposit_state(ctx, q, B.offs, false, out);
  // E = E0 - grad(phi) at every node: lambda_E0 = sum of node lambda_E
  launch_fill<double>(ctx, nc, t->wrk, 1.0);
  launch_mdot<double>(ctx, 3, t->lEn + nc, nc, t->wrk, nc, ctx->red_out);
  double s[3];
  TRY(fetch_n(ctx, ctx->red_out, 3, s));
  for (int a = 0; a < 3; a++) lE0[a] += s[a];
  launch_efield_adj(ctx, out[1], out[2], out[3], t->lphi);
  return CHY_OK;
}

// One forward step (push, sort, deposit, forcing, solve) on the context.
int run_step(chy_ctx *ctx)
{
  chy_step_stats st;
  return ctx_step(ctx, &st);
}

int check_supported(chy_ctx *ctx)
{
  if (ctx->mixed) return set_err(ctx, CHY_ERR_SUP, "the adjoint runs on FP64 contexts (-precision fp64)");
  if (ctx->prm.lhs_type != CHY_LHS_FULL || ctx->prm.rhs_type != CHY_RHS_FULL)
    return set_err(ctx, CHY_ERR_SUP, "the adjoint needs --lhs-type full and --rhs-type full");
  if (ctx->np <= 0) return set_err(ctx, CHY_ERR_ARG_WRONG, "no ions");
  return CHY_OK;
}

}  // namespace

void adjoint_release(chy_ctx *ctx)
{
  if (ctx && ctx->adj) {
    (void)hipStreamSynchronize(ctx->stream);
    free_tape((Tape *)ctx->adj);
    ctx->adj = nullptr;
  }
}

}  // namespace chy

using namespace chy;

extern "C" {

int chy_adjoint_forward(chy_ctx *ctx, int64_t nsteps, int64_t ckpt, double *J)
{
  if (!ctx || nsteps < 0 || ckpt < 1) return CHY_ERR_ARG_WRONG;
  TRY(check_supported(ctx));
  adjoint_release(ctx);
  Tape *t = new Tape();
  ctx->adj = t;
  const int64_t np = ctx->np, nc = ctx->g.nc;
  t->nsteps = nsteps;
  t->ckpt = std::min<int64_t>(ckpt, std::max<int64_t>(nsteps, 1));
  t->np = np;
  const int64_t nck = std::max<int64_t>(1, (nsteps + t->ckpt - 1) / t->ckpt);
#define A(x) do { int r_ = (x); if (r_) { adjoint_release(ctx); return r_; } } while (0)
  // checkpoints and the end state now; the segment and adjoint buffers are
  // allocated by the first backward sweep
  t->ck.resize(nck);
  for (Stage &s : t->ck) A(alloc_stage(ctx, s, np));
  A(alloc_stage(ctx, t->fin, np));
  A(grid_alloc(ctx, t, &t->ln, nc));
  A(grid_alloc(ctx, t, &t->lg, 3 * nc));
  A(grid_alloc(ctx, t, &t->lphi, nc));
  A(grid_alloc(ctx, t, &t->g3, 3 * nc));
  A(grid_alloc(ctx, t, &t->lEn, 4 * nc));
  A(grid_alloc(ctx, t, &t->wrk, nc));
  // moments and potential of the initial ions (phi from a zero guess)
  A(particles_deposit(ctx));
  launch_fill<double>(ctx, nc, ctx->phi, 0.0);
  ctx->phi_valid = true;
  A(ctx_rhs(ctx));
  {
    chy_solve_stats st;
    A(ctx_solve(ctx, &st));
  }
  for (int64_t k = 0; k < nsteps; k++) {
    if (k % t->ckpt == 0) A(save_stage(ctx, t->ck[k / t->ckpt]));
    A(run_step(ctx));
  }
  if (nsteps == 0) A(save_stage(ctx, t->ck[0]));
  A(save_stage(ctx, t->fin));
  double Jv = 0.0;
  A(wave_energy(ctx, t, &Jv));
#undef A
  t->done_forward = true;
  if (J) *J = Jv;
  return chy_sync(ctx);
}

int chy_adjoint_backward(chy_ctx *ctx, double dJ_dE0[3], double *dJ_ddelta)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  Tape *t = (Tape *)ctx->adj;
  if (!t || !t->done_forward) return set_err(ctx, CHY_ERR_ARG_WRONG, "run chy_adjoint_forward first");
  if (ctx->np != t->np) return set_err(ctx, CHY_ERR_ARG_INCOMP, "the ion count changed since the forward run");
  const int64_t N = t->nsteps, K = t->ckpt, np = t->np, nc = ctx->g.nc;
  if (t->seg.empty()) {
    t->seg.resize(K);
    t->perm.assign(K, nullptr);
    int rc = CHY_OK;
    for (Stage &s : t->seg) if (!rc) rc = alloc_stage(ctx, s, np);
    for (uint32_t *&p : t->perm) if (!rc) rc = dalloc(ctx, &p, np);
    for (int a = 0; a < 2 && !rc; a++) rc = alloc_adp(ctx, t, t->lam[a], np);
    for (int a = 0; a < 3 && !rc; a++) rc = grid_alloc(ctx, t, &t->lE[a], np);
    if (rc) { adjoint_release(ctx); return rc; }
  }
  double lE0[3] = {0.0, 0.0, 0.0};
  int cur = 0;  // lam[cur]: adjoint of the ions of the latest stage processed
  for (int a = 0; a < 2; a++) {
    double *f[7] = {t->lam[a].x, t->lam[a].y, t->lam[a].z, t->lam[a].vx, t->lam[a].vy, t->lam[a].vz, t->lam[a].w};
    for (double *p : f) launch_fill<double>(ctx, np, p, 0.0);
  }
  bool have_lphi = false;  // lambda_phi of the stage being entered (none for the final one)
  const int64_t nseg = N > 0 ? (N + K - 1) / K : 0;
  for (int64_t sg = nseg - 1; sg >= 0; sg--) {
    const int64_t c0 = sg * K, c1 = std::min(c0 + K, N);
    // recompute the segment from its checkpoint
    TRY(restore_stage(ctx, t->ck[sg]));
    for (int64_t k = c0; k < c1; k++) {
      ctx->sort_src = t->perm[k - c0];
      const int rc = run_step(ctx);
      ctx->sort_src = nullptr;
      if (rc) return rc;
      TRY(save_stage(ctx, t->seg[k - c0]));
    }
    for (int64_t k = c1 - 1; k >= c0; k--) {
      const Stage &A = t->seg[k - c0];                        // state k + 1
      const Stage &B = k == c0 ? t->ck[sg] : t->seg[k - c0 - 1];  // state k
      TRY(moments_adjoint(ctx, t, A, k + 1 == N, have_lphi ? t->lphi : nullptr, t->lam[cur], lE0));
      TRY(push_adjoint(ctx, t, B, t->perm[k - c0], t->lam[cur], t->lam[1 - cur], lE0));
      cur = 1 - cur;
      have_lphi = true;
    }
  }
  // the initial moments and solve
  const Stage &S0 = t->ck[0];
  TRY(moments_adjoint(ctx, t, S0, N == 0, have_lphi ? t->lphi : nullptr, t->lam[cur], lE0));
  if (dJ_ddelta) {
    hipLaunchKernelGGL(k_cell_sum, nb_for(nc), kBlock, 0, ctx->stream, nc, S0.offs, t->lam[cur].w, ctx->r64);
    CHY_HIP(hipMemcpyAsync(dJ_ddelta, ctx->r64, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
  }
  // leave the context at the end of the forward run
  TRY(restore_stage(ctx, t->fin));
  TRY(particles_deposit(ctx));
  TRY(chy_sync(ctx));
  if (dJ_dE0) for (int a = 0; a < 3; a++) dJ_dE0[a] = lE0[a];
  return CHY_OK;
}

}  // extern "C"
