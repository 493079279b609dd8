// chy_ctx.hip -- context, C ABI entry points and the potential solver.
//
// The solver replaces the PETSc KSP driven at reference hybrid.c:73-78 and
// potential.c:6-55.  It is an outer FP64 defect correction
//     r = b - A(n) phi (FP64),  A e = r / |r| (inner Krylov),  phi += |r| e
// around an inner right-preconditioned GMRES(m) (two-pass classical
// Gram-Schmidt with fused multi-dot kernels) or BiCGStab, in FP64 (plain
// mode, one outer sweep) or FP32 (mixed mode).  Inner products and norms
// always accumulate in FP64.  The constant null space is handled by the
// mean-free forcing, the mean-free FFT preconditioner and a final mean
// projection of phi.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "chy_internal.h"

using namespace chy;

namespace chy {

static bool timer_on(const chy_ctx *ctx, const char *name)
{
  if (!ctx->timing) return false;
  if (ctx->timing_only.empty()) return true;
  for (const std::string &s : ctx->timing_only) if (s == name) return true;
  return false;
}

void timer_start(chy_ctx *ctx, const char *name)
{
  if (!timer_on(ctx, name)) return;
  Timer &t = ctx->timers[name];
  if (t.used == t.a.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    t.a.push_back(a);
    t.b.push_back(b);
  }
  (void)hipEventRecord(t.a[t.used], ctx->stream);
  t.open = true;
}

void timer_stop(chy_ctx *ctx, const char *name)
{
  if (!timer_on(ctx, name)) return;
  Timer &t = ctx->timers[name];
  if (!t.open) return;
  (void)hipEventRecord(t.b[t.used], ctx->stream);
  t.used++;
  t.open = false;
}

// Accumulate the elapsed times of all recorded pairs (synchronises).
static void timers_resolve(chy_ctx *ctx)
{
  for (auto &kv : ctx->timers) {
    Timer &t = kv.second;
    for (size_t i = 0; i < t.used; i++) {
      (void)hipEventSynchronize(t.b[i]);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, t.a[i], t.b[i]);
      t.total_ms += ms;
      t.count++;
    }
    t.used = 0;
  }
}

int wait_seq(chy_ctx *ctx, uint64_t s)
{
  // The device's store arrives within microseconds of the kernel's end,
  // sooner than a stream synchronisation returns.  A wait longer than any
  // single pass falls back to that synchronisation, which also reports a
  // device error; after it the number must be there.
  const volatile uint64_t *p = ctx->h_status + kSeqWord;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t sp = 1; *p < s; sp++) {
    __builtin_ia32_pause();
    if ((sp & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
      CHY_HIP(hipStreamSynchronize(ctx->stream));
      if (*p < s) return set_err(ctx, CHY_ERR_LIB, "read-back sequence number missing after the stream drained");
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return CHY_OK;
}

}  // namespace chy

namespace {

template <typename T> T *mom(chy_ctx *c, int a) { return (T *)c->mom[a]; }

int sync(chy_ctx *ctx)
{
  CHY_HIP(hipStreamSynchronize(ctx->stream));
  return CHY_OK;
}

// Basis slots: GMRES uses V_0..V_m, BiCGStab slots 0..7; the last slot holds
// the inner correction e.
int64_t nvec_V(const chy_ctx *ctx) { return std::max<int64_t>(ctx->restart + 2, 10); }

// ---- preconditioner ------------------------------------------------------------------
template <typename T>
int pc_apply(chy_ctx *ctx, const double v19[19], const T *r, T *z)
{
  const int64_t nc = ctx->g.nc;
  timer_start(ctx, "pc");
  int rc = CHY_OK;
  const int pc = ctx->prm.pc_type;
  T *sb = (T *)ctx->pc_buf;
  const bool fftlike = pc == CHY_PC_FFT || pc == CHY_PC_FFT2 || pc == CHY_PC_FFT2Z;
  if (pc == CHY_PC_MG && ctx->prm.lhs_type != CHY_LHS_IDENTITY && !ctx->solve_transpose) {
    rc = mg_apply<T>(ctx, r, z);
  } else if (pc == CHY_PC_MG && ctx->solve_transpose) {
    // the z-line smoother is built for A(n); transposed solves (the adjoint)
    // take the FFT preconditioner of the mean density, symmetric, or Jacobi
    if (ctx->fft_ok) rc = fft_pc_apply<T>(ctx, v19, r, z);
    else launch_jacobi<T>(ctx, ctx->prm.lhs_type, mom<T>(ctx, 0), r, z);
  } else if (fftlike && ctx->fft_ok && ctx->prm.lhs_type != CHY_LHS_IDENTITY) {
    if (pc == CHY_PC_FFT) {
      rc = fft_pc_apply<T>(ctx, v19, r, z);
    } else {
      // fft2z: z += zline(r - A z) (built for A(n); transposed solves keep fft2)
      const bool zl = pc == CHY_PC_FFT2Z && !ctx->solve_transpose && ctx->zl_buf;
      bool zl_done = false;
      rc = fft2_pc_apply<T>(ctx, v19, r, z, sb, zl, &zl_done);
      if (rc == CHY_OK && zl && !zl_done) {
        launch_residual_t<T>(ctx, ctx->prm.lhs_type, mom<T>(ctx, 0), r, z, sb);
        launch_zline_apply<T>(ctx, (const T *)ctx->zl_buf, ctx->zl_col, sb, z);
      }
    }
  } else if (pc == CHY_PC_ZLINE || (fftlike && !ctx->fft_ok)) {
    // grids the FFT does not take fall back to the z-line blocks
    launch_zline<T>(ctx, ctx->prm.lhs_type, mom<T>(ctx, 0), r, z, sb, sb + nc);
  } else if (pc == CHY_PC_CHEBY) {
    launch_cheby<T>(ctx, ctx->prm.lhs_type, ctx->solve_transpose, ctx->prm.pc_cheby_degree, mom<T>(ctx, 0), r, z, sb,
                    sb + nc, sb + 2 * nc);
  } else if (pc == CHY_PC_JACOBI) {
    launch_jacobi<T>(ctx, ctx->prm.lhs_type, mom<T>(ctx, 0), r, z);
  } else {
    CHY_HIP(hipMemcpyAsync(z, r, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  }
  timer_stop(ctx, "pc");
  return rc;
}

// Read nv doubles from the device reduction buffer.  Values in red_out are
// in its host mirror once the last mirrored combine has stored its number.
int fetch(chy_ctx *ctx, const double *dev, int nv, double *host)
{
  if (ctx->red_mirror && dev >= ctx->red_out && dev + nv <= ctx->red_out + ctx->red_out_n) {
    TRY(wait_seq(ctx, ctx->seq_issued));
    const double *src = ctx->h_red + (dev - ctx->red_out);
    if (src != host) std::memcpy(host, src, sizeof(double) * nv);
    return CHY_OK;
  }
  CHY_HIP(hipMemcpyAsync(host, dev, sizeof(double) * nv, hipMemcpyDeviceToHost, ctx->stream));
  CHY_HIP(hipStreamSynchronize(ctx->stream));
  return CHY_OK;
}

// ---- inner GMRES on A e = rhs (e starts at zero), precision T -------------------
// rhs == V (basis slot 0): the caller wrote the right-hand side there, and
// beta_hint > 0 is its norm; the first cycle then starts without a copy or a
// norm pass, and rhs is saved to `save` before a restart overwrites slot 0.
template <typename T>
int gmres_inner(chy_ctx *ctx, const double v19[19], const T *rhs, T *e, double rtol, int64_t max_it,
                int64_t *its, double *relres, double beta_hint = -1.0, T *save = nullptr)
{
  const int64_t nc = ctx->g.nc, m = ctx->restart;
  T *V = (T *)ctx->V, *Z = (T *)ctx->Z, *w = (T *)ctx->wv;
  const T *n = mom<T>(ctx, 0);
  const int lt = ctx->prm.lhs_type;
  double *hdev = ctx->red_out, *h2dev = ctx->red_out + (m + 2), *ndev = ctx->red_out + 2 * (m + 2);
  std::vector<double> H((m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), h(m + 2), h2(m + 2);
  // The basis is stored unnormalised: V_q = sc[q] U_q with U_q in the V
  // slots, so no pass is spent scaling a new vector.  The dot products come
  // back multiplied by sc[q]^2 (k_combine's scale), which makes them the
  // coefficients of the U_q in the update; the host turns them into the
  // Hessenberg entries.  The preconditioned directions follow:
  // Z_q = sc[q] (M^-1 U_q).
  std::vector<double> sc(m + 2, 0.0);
  double *sc2 = ctx->h_scale;  // sc[q]^2, passed by value to the combines
  launch_fill<T>(ctx, nc, e, (T)0);
  double beta0 = -1.0, beta_prev = -1.0;
  *its = 0;
  *relres = 1.0;
  const bool in_place = rhs == V;
  for (;;) {
    // V0 = rhs - A e
    double nn;
    if (*its == 0) {
      if (!in_place) CHY_HIP(hipMemcpyAsync(V, rhs, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      if (in_place && rhs == V) {  // keep the right-hand side before slot 0 changes
        if (!save) return set_err(ctx, CHY_ERR_LIB, "gmres: no buffer to keep the right-hand side for a restart");
        CHY_HIP(hipMemcpyAsync(save, V, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
        rhs = save;
      }
      if (ctx->solve_transpose) {
        launch_lhs_apply<T, T, T>(ctx, lt, true, n, e, w);
        CHY_HIP(hipMemcpyAsync(V, rhs, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
        launch_axpby<T>(ctx, nc, -1.0, w, 1.0, V);
      } else {  // one stencil pass: V0 and its squared norm's partials
        const int64_t nb = launch_residual_sq<T>(ctx, lt, n, rhs, e, V, ctx->rpl);
        launch_combine_sum(ctx, ctx->rpl, (int)nb, 1, ndev);
      }
    }
    if (*its == 0 && beta_hint > 0.0) {
      nn = beta_hint * beta_hint;
    } else {
      if (*its == 0 || ctx->solve_transpose) launch_mdot<T>(ctx, 1, V, nc, V, nc, ndev);
      TRY(fetch(ctx, ndev, 1, &nn));
    }
    const double beta = sqrt(nn);
    if (beta0 < 0) beta0 = beta;
    if (!(beta0 > 0.0)) { *relres = 0.0; return CHY_OK; }
    *relres = beta / beta0;
    if (!std::isfinite(beta)) return set_err(ctx, CHY_ERR_FP, "non-finite residual in the Krylov solve");
    if (beta <= rtol * beta0 || *its >= max_it) return CHY_OK;
    // FP32: a whole restart cycle that cut the true residual by less than 5 %
    // has hit the rounding floor (about eps_32 times the conditioning): hand
    // back to the FP64 outer sweep
    if (sizeof(T) == 4 && beta_prev > 0.0 && beta > 0.95 * beta_prev) return CHY_OK;
    beta_prev = beta;
    sc[0] = 1.0 / beta;
    sc2[0] = sc[0] * sc[0];
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int64_t j = 0;
    bool done = false;
    for (; j < m && *its < max_it; j++) {
      T *Vj = V + j * nc, *Zj = Z + j * nc, *Vn = V + (j + 1) * nc;
      TRY(pc_apply<T>(ctx, v19, Vj, Zj));
      timer_start(ctx, "lhs_apply");
      launch_lhs_apply<T, T, T>(ctx, lt, ctx->solve_transpose, n, Zj, Vn);  // w built in place in slot j+1
      timer_stop(ctx, "lhs_apply");
      timer_start(ctx, "krylov");
      // classical Gram-Schmidt; the norm of the result is fused into the last
      // update and everything is read back at once.  REFINE_ALWAYS makes a
      // second pass (the preconditioned operator is close to the identity, so
      // one pass can cancel heavily); REFINE_IFNEEDED makes it when the first
      // pass removed more than half of the squared norm (sum h^2 > |w'|^2,
      // with |w|^2 = |w'|^2 + sum h^2)
      const int refine = ctx->prm.ksp_gmres_cgs_refinement;
      const double sj = sc[j];
      double wn;  // |U_{j+1}| before scaling
      launch_gs_dots<T>(ctx, (int)(j + 1), V, nc, Vn, nc, hdev, sc2);
      if (refine == CHY_CGS_REFINE_ALWAYS) {
        launch_gs_update_dots<T>(ctx, (int)(j + 1), V, nc, hdev, Vn, nc, h2dev, sc2);
        launch_gs_update_norm<T>(ctx, (int)(j + 1), V, nc, h2dev, Vn, nc, ndev);
        TRY(fetch(ctx, ctx->red_out, (int)(2 * (m + 2) + 1), ctx->h_red));
        for (int64_t q = 0; q <= j; q++) h[q] = sc[q] > 0.0 ? (ctx->h_red[q] + ctx->h_red[(m + 2) + q]) * sj / sc[q] : 0.0;
        wn = sqrt(ctx->h_red[2 * (m + 2)]);
      } else {
        launch_gs_update_norm<T>(ctx, (int)(j + 1), V, nc, hdev, Vn, nc, ndev);
        TRY(fetch(ctx, ctx->red_out, (int)(2 * (m + 2) + 1), ctx->h_red));
        double proj = 0.0;
        for (int64_t q = 0; q <= j; q++) {
          h[q] = sc[q] > 0.0 ? ctx->h_red[q] * sj / sc[q] : 0.0;
          proj += h[q] * h[q];
        }
        wn = sqrt(ctx->h_red[2 * (m + 2)]);
        if (refine == CHY_CGS_REFINE_IFNEEDED && proj > wn * sj * wn * sj) {
          launch_gs_dots<T>(ctx, (int)(j + 1), V, nc, Vn, nc, h2dev, sc2);
          launch_gs_update_norm<T>(ctx, (int)(j + 1), V, nc, h2dev, Vn, nc, ndev);
          TRY(fetch(ctx, ctx->red_out, (int)(2 * (m + 2) + 1), ctx->h_red));
          for (int64_t q = 0; q <= j; q++) h[q] += sc[q] > 0.0 ? ctx->h_red[(m + 2) + q] * sj / sc[q] : 0.0;
          wn = sqrt(ctx->h_red[2 * (m + 2)]);
        }
      }
      timer_stop(ctx, "krylov");
      h[j + 1] = wn * sj;
      if (!std::isfinite(h[j + 1])) return set_err(ctx, CHY_ERR_FP, "non-finite Arnoldi vector");
      sc[j + 1] = wn > 0.0 ? 1.0 / wn : 0.0;
      sc2[j + 1] = sc[j + 1] * sc[j + 1];
      for (int64_t q = 0; q < j; q++) {
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
      for (int64_t q = 0; q <= j; q++) H[q * m + j] = h[q];
      (*its)++;
      if (ctx->prm.ksp_monitor) printf("  %lld KSP Residual norm %.12e\n", (long long)*its, fabs(g[j + 1]));
      if (fabs(g[j + 1]) <= rtol * beta0) { j++; done = true; break; }
    }
    // e += Z y
    for (int64_t q = j - 1; q >= 0; q--) {
      double s = g[q];
      for (int64_t r = q + 1; r < j; r++) s -= H[q * m + r] * y[r];
      y[q] = s / H[q * m + q];
    }
    // (the coefficients travel by value with the launch)
    for (int64_t q = 0; q < j; q++) y[q] = -y[q] * sc[q];
    launch_mgs_update_host<T>(ctx, (int)j, Z, nc, y.data(), e, nc);
    (void)done;
    // the true residual of the new e is checked at the top of the loop (a
    // cycle that ends on GMRES's own estimate without it leaves phi measurably
    // further from the exact solution: 2.1e-10 against 7.9e-11 at 256^3)
  }
}

// ---- inner BiCGStab (right preconditioned) --------------------------------------------
template <typename T>
int bicgstab_inner(chy_ctx *ctx, const double v19[19], const T *rhs, T *x, double rtol, int64_t max_it,
                   int64_t *its, double *relres)
{
  const int64_t nc = ctx->g.nc;
  T *V = (T *)ctx->V;
  T *r = V, *rh = V + nc, *p = V + 2 * nc, *v = V + 3 * nc, *yv = V + 4 * nc, *s = V + 5 * nc, *zv = V + 6 * nc,
    *t = V + 7 * nc;
  const T *n = mom<T>(ctx, 0);
  const int lt = ctx->prm.lhs_type;
  double *d = ctx->red_out;
  launch_fill<T>(ctx, nc, x, (T)0);
  launch_fill<T>(ctx, nc, p, (T)0);
  launch_fill<T>(ctx, nc, v, (T)0);
  CHY_HIP(hipMemcpyAsync(r, rhs, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  CHY_HIP(hipMemcpyAsync(rh, rhs, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  launch_mdot<T>(ctx, 1, r, nc, r, nc, d);
  double nn;
  TRY(fetch(ctx, d, 1, &nn));
  const double b0 = sqrt(nn);
  double rho = 1.0, alpha = 1.0, omega = 1.0, best = 1.0;
  int64_t best_it = 0;
  *its = 0;
  *relres = 1.0;
  if (!(b0 > 0.0)) { *relres = 0.0; return CHY_OK; }
  while (*its < max_it) {
    // FP32: no progress for 50 iterations near the rounding floor
    if (*relres < best) { best = *relres; best_it = *its; }
    else if (sizeof(T) == 4 && *its - best_it > 50 && best < 1e-3) return CHY_OK;
    launch_mdot<T>(ctx, 1, rh, nc, r, nc, d);
    double rho1;
    TRY(fetch(ctx, d, 1, &rho1));
    const double beta = (rho1 / rho) * (alpha / omega);
    // p = r + beta (p - omega v)
    launch_axpby<T>(ctx, nc, -omega, v, 1.0, p);
    launch_axpby<T>(ctx, nc, 1.0, r, beta, p);
    TRY(pc_apply<T>(ctx, v19, p, yv));
    launch_lhs_apply<T, T, T>(ctx, lt, ctx->solve_transpose, n, yv, v);
    launch_mdot<T>(ctx, 1, rh, nc, v, nc, d);
    double rhv;
    TRY(fetch(ctx, d, 1, &rhv));
    alpha = rho1 / rhv;
    CHY_HIP(hipMemcpyAsync(s, r, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
    launch_axpby<T>(ctx, nc, -alpha, v, 1.0, s);
    launch_axpby<T>(ctx, nc, alpha, yv, 1.0, x);
    (*its)++;
    launch_mdot<T>(ctx, 1, s, nc, s, nc, d);
    double ss;
    TRY(fetch(ctx, d, 1, &ss));
    *relres = sqrt(ss) / b0;
    if (!std::isfinite(*relres)) return set_err(ctx, CHY_ERR_FP, "non-finite residual in BiCGStab");
    if (*relres <= rtol) return CHY_OK;
    TRY(pc_apply<T>(ctx, v19, s, zv));
    launch_lhs_apply<T, T, T>(ctx, lt, ctx->solve_transpose, n, zv, t);
    launch_mdot<T>(ctx, 1, t, nc, s, nc, d);
    launch_mdot<T>(ctx, 1, t, nc, t, nc, d + 1);
    double ts[2];
    TRY(fetch(ctx, d, 2, ts));
    omega = ts[1] > 0.0 ? ts[0] / ts[1] : 0.0;
    launch_axpby<T>(ctx, nc, omega, zv, 1.0, x);
    CHY_HIP(hipMemcpyAsync(r, s, sizeof(T) * nc, hipMemcpyDeviceToDevice, ctx->stream));
    launch_axpby<T>(ctx, nc, -omega, t, 1.0, r);
    launch_mdot<T>(ctx, 1, r, nc, r, nc, d);
    double rr;
    TRY(fetch(ctx, d, 1, &rr));
    *relres = sqrt(rr) / b0;
    if (*relres <= rtol) return CHY_OK;
    rho = rho1;
  }
  return CHY_OK;
}

// Row of the constant-coefficient operator at the mean density (the FFT
// preconditioner's symbol).
// With bb given, |b|^2 is reduced beside the density sum and read back with it
// (one host round trip for both).
template <typename T>
int pc_symbol(chy_ctx *ctx, double v19[19], double *bb = nullptr)
{
  const int64_t nc = ctx->g.nc;
  launch_sum<T>(ctx, mom<T>(ctx, 0), nc, ctx->red_out);
  if (bb) launch_mdot<double>(ctx, 1, ctx->b, nc, ctx->b, nc, ctx->red_out + 1);
  double sv[2] = {0.0, 0.0};
  TRY(fetch(ctx, ctx->red_out, bb ? 2 : 1, sv));
  if (bb) *bb = sv[1];
  uniform_row(ctx->prm.lhs_type, ctx->prm.d, ctx->prm.kappa_e, sv[0] / (double)nc, v19);
  // coefficients of the z-averaged operator for the two-level preconditioner
  const bool two = ctx->prm.pc_type == CHY_PC_FFT2 || ctx->prm.pc_type == CHY_PC_FFT2Z;
  // and the z-line factors of the fft2z correction (for the full operator,
  // both from one walk up the columns)
  const bool coef = two && ctx->fft_ok && ctx->pc2_buf;
  const bool fact = ctx->prm.pc_type == CHY_PC_FFT2Z && ctx->fft_ok && ctx->zl_buf;
  if (ctx->prm.pc_type == CHY_PC_MG && ctx->mg && ctx->prm.lhs_type != CHY_LHS_IDENTITY)
    TRY(mg_prepare<T>(ctx, mom<T>(ctx, 0)));
  if (ctx->prm.lhs_type == CHY_LHS_FULL && (coef || fact)) {
    launch_zcolumn<T>(ctx, mom<T>(ctx, 0), coef ? (T *)ctx->pc2_buf : nullptr, fact ? (T *)ctx->zl_buf : nullptr,
                      fact ? ctx->zl_col : nullptr);
  } else {
    if (coef) launch_zcoef<T>(ctx, ctx->prm.lhs_type, mom<T>(ctx, 0), (T *)ctx->pc2_buf);
    if (fact) launch_zline_factor<T>(ctx, ctx->prm.lhs_type, mom<T>(ctx, 0), (T *)ctx->zl_buf, ctx->zl_col);
  }
  return CHY_OK;
}

// Outer FP64 defect correction; phi (device, FP64) in/out, b (device) mean-free.
template <typename T>
int solve_impl(chy_ctx *ctx, chy_solve_stats *st)
{
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t nc = ctx->g.nc;
  const chy_params &p = ctx->prm;
  const T *n = mom<T>(ctx, 0);
  double v19[19], bb = 0.0;
  TRY(pc_symbol<T>(ctx, v19, &bb));
  const double bn = sqrt(bb);
  const double tol = std::max(p.ksp_rtol * bn, p.ksp_atol);
  int64_t its = 0, outer = 0;
  double rn = 0.0;
  int converged = 0;
  T *rhsT = (T *)ctx->t64;  // FP64-sized buffer holds the inner right-hand side
  const bool gm_in_place = !ctx->solve_transpose && p.ksp_type != CHY_KSP_BCGS;
  T *ev = (T *)ctx->V + (nvec_V(ctx) - 1) * nc;
  for (;;) {
    // r = b - A phi in FP64 arithmetic, stored in the inner precision as the
    // inner solve's right-hand side (unscaled), and |r|^2 in FP64
    if (ctx->solve_transpose) {
      launch_lhs_apply<T, double, double>(ctx, p.lhs_type, true, n, ctx->phi, ctx->r64);
      launch_axpby<double>(ctx, nc, 1.0, ctx->b, -1.0, ctx->r64);
      launch_mdot<double>(ctx, 1, ctx->r64, nc, ctx->r64, nc, ctx->red_out);
      launch_convert<double, T>(ctx, nc, ctx->r64, 1.0, rhsT);
    } else {
      // GMRES takes it in its first basis slot (no copy); BiCGStab in rhsT
      launch_residual_norm<T, T>(ctx, p.lhs_type, n, ctx->b, ctx->phi, gm_in_place ? (T *)ctx->V : rhsT, ctx->rpl);
      launch_combine_sum(ctx, ctx->rpl, (int)lhs_zsum_blocks(ctx->g, false), 1, ctx->red_out);
    }
    double rr;
    TRY(fetch(ctx, ctx->red_out, 1, &rr));
    rn = sqrt(rr);
    if (p.ksp_monitor)
      printf("  outer %lld: true residual %.6e (relative %.3e), inner iterations so far %lld\n", (long long)outer, rn,
             bn > 0 ? rn / bn : rn, (long long)its);
    if (!std::isfinite(rn)) return set_err(ctx, CHY_ERR_FP, "non-finite residual in the potential solve");
    if (rn <= tol) { converged = 1; break; }
    if (its >= p.ksp_max_it || outer >= 64) break;
    // FP32 inner solves stop at a floor above their rounding limit (about
    // eps_32 times the preconditioned operator's conditioning); FP64 solves go
    // to tol
    const double inner_rtol =
        ctx->mixed ? std::max(p.ksp_inner_rtol, 0.5 * tol / rn) : std::min(1.0, 0.5 * tol / rn);
    int64_t it_in = 0;
    double rel = 1.0;
    if (p.ksp_type == CHY_KSP_BCGS)
      TRY(bicgstab_inner<T>(ctx, v19, rhsT, ev, inner_rtol, p.ksp_max_it - its, &it_in, &rel));
    else
      TRY(gm_in_place ? gmres_inner<T>(ctx, v19, (T *)ctx->V, ev, inner_rtol, p.ksp_max_it - its, &it_in, &rel, rn, rhsT)
                      : gmres_inner<T>(ctx, v19, rhsT, ev, inner_rtol, p.ksp_max_it - its, &it_in, &rel));
    its += it_in;
    outer++;
    // phi += e
    launch_add_scaled<T>(ctx, nc, ev, 1.0, ctx->phi);
    if (it_in == 0) break;
  }
  // mean projection of phi (null space)
  launch_sum<double>(ctx, ctx->phi, nc, ctx->red_out);
  launch_shift<double>(ctx, nc, ctx->phi, ctx->red_out, -1.0 / (double)nc);
  TRY(sync(ctx));
  const auto t1 = std::chrono::steady_clock::now();
  if (st) {
    st->iterations = its;
    st->outer = outer;
    st->residual = bn > 0.0 ? rn / bn : rn;
    st->converged = converged;
    st->reason = converged ? (rn <= p.ksp_rtol * bn ? 2 : 3) : -3;
    st->ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  if (p.ksp_converged_reason)
    printf("Linear solve %s due to %s iterations %lld\n", converged ? "converged" : "did not converge",
           converged ? "CONVERGED_RTOL" : "DIVERGED_ITS", (long long)its);
  if (!converged && p.ksp_error_if_not_converged)
    return set_err(ctx, CHY_ERR_NOT_CONVERGED, "KSPSolve has not converged");
  ctx->phi_valid = true;
  return CHY_OK;
}

int do_rhs(chy_ctx *ctx)
{
  if (!ctx->moments_valid) return set_err(ctx, CHY_ERR_ARG_WRONG, "no moments: collect or set them first");
  timer_start(ctx, "rhs");
  if (ctx->mixed)
    launch_rhs<float>(ctx, ctx->prm.rhs_type, mom<float>(ctx, 0), mom<float>(ctx, 1), mom<float>(ctx, 2),
                      mom<float>(ctx, 3), ctx->b);
  else
    launch_rhs<double>(ctx, ctx->prm.rhs_type, mom<double>(ctx, 0), mom<double>(ctx, 1), mom<double>(ctx, 2),
                       mom<double>(ctx, 3), ctx->b);
  timer_stop(ctx, "rhs");
  return CHY_OK;
}

int do_solve(chy_ctx *ctx, chy_solve_stats *st)
{
  if (!ctx->phi_valid) launch_fill<double>(ctx, ctx->g.nc, ctx->phi, 0.0);  // potential.c:20-29
  return ctx->mixed ? solve_impl<float>(ctx, st) : solve_impl<double>(ctx, st);
}

int do_efield(chy_ctx *ctx)
{
  if (!ctx->phi_valid) launch_fill<double>(ctx, ctx->g.nc, ctx->phi, 0.0);
  if (ctx->mixed) launch_efield<float>(ctx, ctx->phi, (float *)ctx->efield[0]);
  else launch_efield<double>(ctx, ctx->phi, (double *)ctx->efield[0]);
  return CHY_OK;
}

bool check(const chy_ctx *c) { return c != nullptr; }

}  // namespace

// =============================== C ABI ===================================
extern "C" {

int chy_create(const chy_params *prm_in, int64_t np_capacity, chy_ctx **out)
{
  if (!prm_in || !out) return CHY_ERR_ARG_WRONG;
  *out = nullptr;
  chy_ctx *ctx = new chy_ctx();
  ctx->prm = *prm_in;
  char err[2048] = "";
  int rc = chy_params_derive(&ctx->prm, err, sizeof err);
  if (rc) { int r = rc; delete ctx; fprintf(stderr, "chy_create: %s", err); return r; }
  const chy_params &p = ctx->prm;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete ctx;
    return CHY_ERR_LIB;
  }
  if (hipSetDevice(0) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) { delete ctx; return CHY_ERR_LIB; }
  if (hipEventCreate(&ctx->ev_step) != hipSuccess || hipEventCreate(&ctx->ev_particles) != hipSuccess) {
    chy_destroy(ctx);
    return CHY_ERR_LIB;
  }
  ctx->mixed = p.precision == CHY_PREC_MIXED;
  Grid &g = ctx->g;
  g.nx = (int)p.N[0]; g.ny = (int)p.N[1]; g.nz = (int)p.N[2];
  g.mx = fastdiv_magic((uint32_t)g.nx);
  g.my = fastdiv_magic((uint32_t)g.ny);
  g.nc = p.N[0] * p.N[1] * p.N[2];
  for (int a = 0; a < 3; a++) { g.p0[a] = p.p0[a]; g.p1[a] = p.p1[a]; g.d[a] = p.d[a]; g.L[a] = p.L[a]; }
  // node counts index 32-bit; row launches put ny and nz on the y and z grid axes
  if (g.nc >= (int64_t)1 << 31 || g.ny > 65535 || g.nz > 65535) { delete ctx; return CHY_ERR_ARG_RANGE; }
  ctx->cap = np_capacity > 0 ? np_capacity : std::max<int64_t>(p.Np, 1);
  // slots, offsets and sort permutations are 32-bit
  if (ctx->cap > (int64_t)0xFFFFFFFF) { delete ctx; return CHY_ERR_ARG_RANGE; }
  ctx->weighted = p.perturbation != CHY_PERT_NONE || p.weighted;
  ctx->restart = std::max<int64_t>(1, std::min<int64_t>(p.ksp_restart, 64));
  const int64_t nc = g.nc, cap = ctx->cap;
  const size_t ts = ctx->mixed ? 4 : 8;
#define A(x) do { int r_ = (x); if (r_) { chy_destroy(ctx); return r_; } } while (0)
  for (int s = 0; s < 2; s++) {
    if (ctx->mixed) A(alloc_state<float>(ctx, ctx->pf[s], cap, ctx->weighted));
    else A(alloc_state<double>(ctx, ctx->pd[s], cap, ctx->weighted));
  }
  A(dalloc(ctx, &ctx->perm, cap));
  A(dalloc(ctx, &ctx->slotkey, cap));
  A(dalloc(ctx, &ctx->counts, nc + 1));
  A(dalloc(ctx, &ctx->offsets, nc + 1));
  A(dalloc(ctx, &ctx->cursor, nc + 1));
  A(dalloc(ctx, &ctx->scan_tmp, 2 * (nc / 4096 + 2)));  // tile sums (+ total), then tile maxima
  A(dalloc(ctx, &ctx->maxcount, 1));
  A(dalloc(ctx, &ctx->trans, 27 * nc));
  A(dalloc(ctx, &ctx->info, cap));
  A(dalloc(ctx, &ctx->dir, cap + 16));  // the push's move directions; read as aligned words
  A(dalloc(ctx, &ctx->nbase, 27 * nc));
  A(dalloc(ctx, &ctx->offs_alt, nc + 1));
  A(dalloc(ctx, &ctx->sortflag, 2));
  launch_fill<uint32_t>(ctx, 2, ctx->sortflag, 0u);  // [1]: placements lost, sticky until read
  {
    unsigned char *pp = nullptr;
    A(dalloc(ctx, &pp, (int64_t)(32 * ts) * nc));
    ctx->partials = pp;
  }
  A(dalloc(ctx, &ctx->sred, std::max<int64_t>(4096, nc / 32 + 1)));  // normaliser partials, one per block
  A(dalloc(ctx, &ctx->S_dev, 1));
  A(dalloc(ctx, &ctx->cstats, 6));
  A(dalloc(ctx, &ctx->push_part, 4 * ((int64_t)kPushBlocks + 64)));
  for (int a = 0; a < 4; a++) { unsigned char *q = nullptr; A(dalloc(ctx, &q, (int64_t)ts * nc)); ctx->mom[a] = q; }
  { unsigned char *q = nullptr; A(dalloc(ctx, &q, (int64_t)ts * 4 * nc)); ctx->efield[0] = q; }
  A(dalloc(ctx, &ctx->phi, nc));
  if (p.ksp_initial_guess == CHY_GUESS_EXTRAPOLATE) A(dalloc(ctx, &ctx->phi_prev, nc));
  A(dalloc(ctx, &ctx->b, nc));
  A(dalloc(ctx, &ctx->r64, nc));
  A(dalloc(ctx, &ctx->t64, nc));
  {
    unsigned char *q = nullptr;
    A(dalloc(ctx, &q, (int64_t)ts * nc * nvec_V(ctx))); ctx->V = q;
    A(dalloc(ctx, &q, (int64_t)ts * nc * ctx->restart)); ctx->Z = q;
    A(dalloc(ctx, &q, (int64_t)ts * nc)); ctx->wv = q;
    A(dalloc(ctx, &q, (int64_t)ts * 2 * (nc + 32 * (int64_t)ctx->g.nz))); ctx->fft_buf = q;  // half spectrum, padded planes
    A(dalloc(ctx, &q, (int64_t)ts * 3 * nc)); ctx->pc_buf = q;
  }
  A(dalloc(ctx, &ctx->red, (int64_t)kRedBlocks * 32));
  A(dalloc(ctx, &ctx->rpl, lhs_zsum_blocks(g, false)));
  ctx->red_out_n = 3 * (ctx->restart + 2) + 16;
  A(dalloc(ctx, &ctx->red_out, ctx->red_out_n));
  // the mirror is coherent host memory: device writes reach it without a copy
  if (hipHostMalloc((void **)&ctx->h_red, sizeof(double) * ctx->red_out_n,
                    hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&ctx->red_mirror, ctx->h_red, 0) != hipSuccess ||
      hipHostMalloc((void **)&ctx->h_scale, sizeof(double) * (ctx->restart + 2)) != hipSuccess ||
      hipHostMalloc((void **)&ctx->h_status, sizeof(uint64_t) * kStatusWords,
                    hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&ctx->status_dev, ctx->h_status, 0) != hipSuccess) {
    chy_destroy(ctx);
    return CHY_ERR_MEM;
  }
  std::memset(ctx->h_status, 0, sizeof(uint64_t) * kStatusWords);  // the sequence word starts below every number
  if (p.pc_type == CHY_PC_MG) A(mg_setup(ctx));
  A(dalloc(ctx, &ctx->diag_part, diag_elems()));
  ctx->diag_out = ctx->diag_part + diag_parts();
  A(diag_scratch_init(ctx));
  A(fft_setup(ctx));
  if ((p.pc_type == CHY_PC_FFT2 || p.pc_type == CHY_PC_FFT2Z) && ctx->fft_ok) {
    unsigned char *q = nullptr;
    A(dalloc(ctx, &q, (int64_t)ts * fft2_buf_elems(ctx)));
    ctx->pc2_buf = q;
    A(dalloc(ctx, &ctx->pc2_h, fft2_h_elems(ctx)));
    if (p.pc_type == CHY_PC_FFT2Z && ctx->g.nz >= 3) {
      unsigned char *zq = nullptr;
      A(dalloc(ctx, &zq, (int64_t)ts * 4 * nc));
      ctx->zl_buf = zq;
      A(dalloc(ctx, &ctx->zl_col, 2 * (int64_t)ctx->g.nx * ctx->g.ny));
    }
  }
  launch_fill<double>(ctx, nc, ctx->phi, 0.0);
  launch_fill<double>(ctx, 1, ctx->S_dev, 0.0);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) { chy_destroy(ctx); return CHY_ERR_LIB; }
#undef A
  *out = ctx;
  return CHY_OK;
}

static int out_join(chy_ctx *ctx);

int chy_destroy(chy_ctx *ctx)
{
  if (!ctx) return CHY_OK;
  (void)out_join(ctx);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->out_stream) (void)hipStreamSynchronize(ctx->out_stream);
  if (ctx->out_host) (void)hipHostFree(ctx->out_host);
  if (ctx->out_dev) (void)hipFree(ctx->out_dev);
  if (ctx->out_snap) (void)hipEventDestroy(ctx->out_snap);
  if (ctx->out_copied) (void)hipEventDestroy(ctx->out_copied);
  if (ctx->out_stream) (void)hipStreamDestroy(ctx->out_stream);
  adjoint_release(ctx);
  mg_release(ctx);
  for (int s = 0; s < 2; s++) { free_state(ctx->pd[s]); free_state(ctx->pf[s]); }
  void *ps[] = {ctx->perm, ctx->slotkey, ctx->counts, ctx->offsets, ctx->cursor, ctx->scan_tmp, ctx->maxcount,
                ctx->trans, ctx->info, ctx->dir, ctx->nbase, ctx->offs_alt, ctx->sortflag,
                ctx->partials,
                ctx->sred, ctx->S_dev, ctx->cstats, ctx->push_part, ctx->mom[0], ctx->mom[1], ctx->mom[2], ctx->mom[3],
                ctx->efield[0], ctx->phi, ctx->phi_prev, ctx->b, ctx->r64, ctx->t64, ctx->V,
                ctx->Z, ctx->wv, ctx->fft_buf, ctx->pc_buf, ctx->pc2_buf, ctx->pc2_h, ctx->zl_buf, ctx->zl_col, ctx->red, ctx->rpl, ctx->red_out, ctx->diag_part, ctx->tw[0], ctx->tw[1], ctx->tw[2],
                ctx->trig[0], ctx->trig[1], ctx->trig[2]};
  for (void *p : ps) if (p) (void)hipFree(p);
  if (ctx->h_red) (void)hipHostFree(ctx->h_red);
  if (ctx->h_scale) (void)hipHostFree(ctx->h_scale);
  if (ctx->h_status) (void)hipHostFree(ctx->h_status);
  for (auto &kv : ctx->timers) {
    for (hipEvent_t e : kv.second.a) (void)hipEventDestroy(e);
    for (hipEvent_t e : kv.second.b) (void)hipEventDestroy(e);
  }
  if (ctx->ev_step) (void)hipEventDestroy(ctx->ev_step);
  if (ctx->ev_particles) (void)hipEventDestroy(ctx->ev_particles);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CHY_OK;
}

const char *chy_last_error(const chy_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int chy_sync(chy_ctx *ctx) { return check(ctx) ? sync(ctx) : CHY_ERR_ARG_WRONG; }

void *chy_stream(chy_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int chy_get_params(const chy_ctx *ctx, chy_params *prm)
{
  if (!ctx || !prm) return CHY_ERR_ARG_WRONG;
  *prm = ctx->prm;
  return CHY_OK;
}

int64_t chy_num_particles(const chy_ctx *ctx) { return ctx ? ctx->np : -1; }

int chy_set_tolerances(chy_ctx *ctx, double rtol, double atol, int64_t max_it)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  if (!(rtol < 1.0) || !(atol == atol)) return CHY_ERR_ARG_RANGE;  // NaN or rtol >= 1
  if (rtol >= 0.0) ctx->prm.ksp_rtol = rtol;
  if (atol >= 0.0) ctx->prm.ksp_atol = atol;
  if (max_it >= 0) ctx->prm.ksp_max_it = max_it;
  return CHY_OK;
}

int chy_set_step(chy_ctx *ctx, int64_t it)
{
  if (!ctx || it < 0) return CHY_ERR_ARG_WRONG;
  ctx->step = it;
  return CHY_OK;
}

int chy_set_particles(chy_ctx *ctx, int64_t np, const double *pos, const double *vel, const double *w,
                      const uint32_t *pid)
{
  if (!ctx || (np > 0 && (!pos || !vel))) return CHY_ERR_ARG_WRONG;
  int rc = particles_upload(ctx, np, pos, vel, w, pid);
  if (rc) return rc;
  return sync(ctx);
}

int chy_get_particles(chy_ctx *ctx, double *pos, double *vel, double *w, uint32_t *pid)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  return particles_download(ctx, pos, vel, w, pid);
}

int chy_init_particles(chy_ctx *ctx)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  int rc = particles_init(ctx);
  if (rc) return rc;
  return sync(ctx);
}

int chy_update_velocities(chy_ctx *ctx, chy_collision_stats *stats)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  do_efield(ctx);
  const int rc = particles_push(ctx, false, ctx->prm.mn > 0.0, stats);
  ctx->step++;
  return rc;
}

int chy_update_positions(chy_ctx *ctx)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  int rc = particles_move_only(ctx);
  if (rc) return rc;
  rc = particles_sort(ctx);
  if (rc) return rc;
  ctx->moments_valid = false;
  return sync(ctx);
}

int chy_collect_moments(chy_ctx *ctx, double *vlasov)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  int rc = particles_deposit(ctx);
  if (rc) return rc;
  if (vlasov) {
    rc = sync(ctx);
    if (rc) return rc;
    return chy_get_moments(ctx, vlasov);
  }
  return sync(ctx);
}

int chy_set_moments(chy_ctx *ctx, const double *vlasov)
{
  if (!ctx || !vlasov) return CHY_ERR_ARG_WRONG;
  const int64_t nc = ctx->g.nc;
  for (int a = 0; a < 4; a++) {
    if (ctx->mixed) {
      std::vector<float> f(nc);
      for (int64_t c = 0; c < nc; c++) f[c] = (float)vlasov[4 * c + a];
      CHY_HIP(hipMemcpy(ctx->mom[a], f.data(), sizeof(float) * nc, hipMemcpyHostToDevice));
    } else {
      std::vector<double> f(nc);
      for (int64_t c = 0; c < nc; c++) f[c] = vlasov[4 * c + a];
      CHY_HIP(hipMemcpy(ctx->mom[a], f.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
    }
  }
  ctx->moments_valid = true;
  return CHY_OK;
}

int chy_get_moments(chy_ctx *ctx, double *vlasov)
{
  if (!ctx || !vlasov) return CHY_ERR_ARG_WRONG;
  const int64_t nc = ctx->g.nc;
  if (!ctx->moments_valid) { for (int64_t c = 0; c < 4 * nc; c++) vlasov[c] = 0.0; return CHY_OK; }
  for (int a = 0; a < 4; a++) {
    std::vector<double> m(nc);
    if (ctx->mixed) {
      std::vector<float> f(nc);
      CHY_HIP(hipMemcpy(f.data(), ctx->mom[a], sizeof(float) * nc, hipMemcpyDeviceToHost));
      for (int64_t c = 0; c < nc; c++) m[c] = f[c];
    } else {
      CHY_HIP(hipMemcpy(m.data(), ctx->mom[a], sizeof(double) * nc, hipMemcpyDeviceToHost));
    }
    for (int64_t c = 0; c < nc; c++) vlasov[4 * c + a] = m[c];
  }
  return CHY_OK;
}

int chy_compute_rhs(chy_ctx *ctx, double *b)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  int rc = do_rhs(ctx);
  if (rc) return rc;
  if (b) CHY_HIP(hipMemcpy(b, ctx->b, sizeof(double) * ctx->g.nc, hipMemcpyDeviceToHost));
  return sync(ctx);
}

static int lhs_apply_host(chy_ctx *ctx, const double *x, double *y, bool transpose)
{
  if (!ctx || !x || !y) return CHY_ERR_ARG_WRONG;
  if (!ctx->moments_valid && ctx->prm.lhs_type == CHY_LHS_FULL)
    return set_err(ctx, CHY_ERR_ARG_WRONG, "no density: collect or set moments first");
  const int64_t nc = ctx->g.nc;
  CHY_HIP(hipMemcpy(ctx->r64, x, sizeof(double) * nc, hipMemcpyHostToDevice));
  timer_start(ctx, "lhs_apply");
  if (ctx->mixed)
    launch_lhs_apply<float, double, double>(ctx, ctx->prm.lhs_type, transpose, mom<float>(ctx, 0), ctx->r64, ctx->t64);
  else
    launch_lhs_apply<double, double, double>(ctx, ctx->prm.lhs_type, transpose, mom<double>(ctx, 0), ctx->r64,
                                             ctx->t64);
  timer_stop(ctx, "lhs_apply");
  CHY_HIP(hipMemcpyAsync(y, ctx->t64, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
  return sync(ctx);
}

int chy_lhs_apply(chy_ctx *ctx, const double *x, double *y) { return lhs_apply_host(ctx, x, y, false); }
int chy_lhs_apply_transpose(chy_ctx *ctx, const double *x, double *y) { return lhs_apply_host(ctx, x, y, true); }

int chy_solve_potential(chy_ctx *ctx, double *phi, chy_solve_stats *st)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = do_rhs(ctx);
  if (rc) return rc;
  rc = do_solve(ctx, st);
  if (st) st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc) return rc;
  if (phi) CHY_HIP(hipMemcpy(phi, ctx->phi, sizeof(double) * ctx->g.nc, hipMemcpyDeviceToHost));
  return CHY_OK;
}

int chy_solve(chy_ctx *ctx, const double *b, double *phi, chy_solve_stats *st)
{
  if (!ctx || !b || !phi) return CHY_ERR_ARG_WRONG;
  if (!ctx->moments_valid && ctx->prm.lhs_type == CHY_LHS_FULL)
    return set_err(ctx, CHY_ERR_ARG_WRONG, "no density: collect or set moments first");
  const int64_t nc = ctx->g.nc;
  CHY_HIP(hipMemcpy(ctx->b, b, sizeof(double) * nc, hipMemcpyHostToDevice));
  CHY_HIP(hipMemcpy(ctx->phi, phi, sizeof(double) * nc, hipMemcpyHostToDevice));
  // The full and Laplacian operators carry the constant null space
  // (lhs.c:334-335,530-531), which the solve projects out of b as PETSc's KSP
  // does for a matrix with an attached null space; a b with a non-zero mean
  // would otherwise be inconsistent
  if (ctx->prm.lhs_type != CHY_LHS_IDENTITY) {
    launch_sum<double>(ctx, ctx->b, nc, ctx->red_out);
    launch_shift<double>(ctx, nc, ctx->b, ctx->red_out, -1.0 / (double)nc);
  }
  ctx->phi_valid = true;
  int rc = do_solve(ctx, st);
  if (rc) return rc;
  CHY_HIP(hipMemcpy(phi, ctx->phi, sizeof(double) * nc, hipMemcpyDeviceToHost));
  return CHY_OK;
}

int chy_pc_apply(chy_ctx *ctx, const double *r, double *z)
{
  if (!ctx || !r || !z) return CHY_ERR_ARG_WRONG;
  if (!ctx->moments_valid && ctx->prm.lhs_type == CHY_LHS_FULL)
    return set_err(ctx, CHY_ERR_ARG_WRONG, "no density: collect or set moments first");
  const int64_t nc = ctx->g.nc;
  CHY_HIP(hipMemcpy(ctx->r64, r, sizeof(double) * nc, hipMemcpyHostToDevice));
  double v19[19];
  if (ctx->mixed) {
    TRY(pc_symbol<float>(ctx, v19));
    float *rT = (float *)ctx->t64, *zT = (float *)ctx->wv;
    launch_convert<double, float>(ctx, nc, ctx->r64, 1.0, rT);
    TRY(pc_apply<float>(ctx, v19, rT, zT));
    launch_convert<float, double>(ctx, nc, zT, 1.0, ctx->t64);
  } else {
    TRY(pc_symbol<double>(ctx, v19));
    TRY(pc_apply<double>(ctx, v19, ctx->r64, ctx->t64));
  }
  CHY_HIP(hipMemcpyAsync(z, ctx->t64, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
  return sync(ctx);
}

int chy_get_potential(chy_ctx *ctx, double *phi)
{
  if (!ctx || !phi) return CHY_ERR_ARG_WRONG;
  if (!ctx->phi_valid) { for (int64_t c = 0; c < ctx->g.nc; c++) phi[c] = 0.0; return CHY_OK; }
  CHY_HIP(hipMemcpy(phi, ctx->phi, sizeof(double) * ctx->g.nc, hipMemcpyDeviceToHost));
  return CHY_OK;
}

int chy_set_potential(chy_ctx *ctx, const double *phi)
{
  if (!ctx || !phi) return CHY_ERR_ARG_WRONG;
  CHY_HIP(hipMemcpy(ctx->phi, phi, sizeof(double) * ctx->g.nc, hipMemcpyHostToDevice));
  ctx->phi_valid = true;
  ctx->phi_prev_valid = false;
  return CHY_OK;
}

// Wait for the asynchronous writer, if one runs; its result.
static int out_join(chy_ctx *ctx)
{
  if (ctx->out_thread.joinable()) ctx->out_thread.join();
  const int rc = ctx->out_rc;
  ctx->out_rc = CHY_OK;
  return rc;
}

int chy_output_write_async(chy_ctx *ctx, const char *stem, const char *insert)
{
  if (!ctx || !stem || !insert) return CHY_ERR_ARG_WRONG;
  const int prev = out_join(ctx);
  if (prev != CHY_OK && prev != CHY_ERR_SUP) return set_err(ctx, prev, "the previous asynchronous output failed");
  const int64_t nc = ctx->g.nc;
  if (!ctx->out_copied) {  // first call: all of the resources, or none
    bool ok = hipMalloc(&ctx->out_dev, sizeof(double) * 6 * nc) == hipSuccess;
    ok = ok && hipHostMalloc(&ctx->out_host, sizeof(double) * 6 * nc) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&ctx->out_stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&ctx->out_snap, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&ctx->out_copied, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      if (ctx->out_snap) (void)hipEventDestroy(ctx->out_snap);
      if (ctx->out_stream) (void)hipStreamDestroy(ctx->out_stream);
      if (ctx->out_host) (void)hipHostFree(ctx->out_host);
      if (ctx->out_dev) (void)hipFree(ctx->out_dev);
      ctx->out_snap = ctx->out_copied = nullptr;
      ctx->out_stream = nullptr;
      ctx->out_host = ctx->out_dev = nullptr;
      (void)hipGetLastError();
      return set_err(ctx, CHY_ERR_MEM, "no memory for the asynchronous output (6 nc f64 on the device and pinned)");
    }
  }
  // snapshot in stream order (the next step's kernels then overwrite the
  // fields freely): moments in f64, rhs, potential; zero where not yet made
  double *d = ctx->out_dev;
  for (int a = 0; a < 4; a++) {
    if (!ctx->moments_valid) launch_fill<double>(ctx, nc, d + a * nc, 0.0);
    else if (ctx->mixed) launch_convert<float, double>(ctx, nc, (const float *)ctx->mom[a], 1.0, d + a * nc);
    else CHY_HIP(hipMemcpyAsync(d + a * nc, ctx->mom[a], sizeof(double) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  }
  CHY_HIP(hipMemcpyAsync(d + 4 * nc, ctx->b, sizeof(double) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  if (ctx->phi_valid)
    CHY_HIP(hipMemcpyAsync(d + 5 * nc, ctx->phi, sizeof(double) * nc, hipMemcpyDeviceToDevice, ctx->stream));
  else
    launch_fill<double>(ctx, nc, d + 5 * nc, 0.0);
  CHY_HIP(hipEventRecord(ctx->out_snap, ctx->stream));
  // the copy to the host overlaps whatever the context stream runs next
  CHY_HIP(hipStreamWaitEvent(ctx->out_stream, ctx->out_snap, 0));
  CHY_HIP(hipMemcpyAsync(ctx->out_host, d, sizeof(double) * 6 * nc, hipMemcpyDeviceToHost, ctx->out_stream));
  CHY_HIP(hipEventRecord(ctx->out_copied, ctx->out_stream));
  const std::string st(stem), in(insert);
  const int dev = ctx->device;
  auto write = [ctx, st, in, dev, nc]() {
    int rc = hipSetDevice(dev) == hipSuccess && hipEventSynchronize(ctx->out_copied) == hipSuccess ? CHY_OK
                                                                                                   : CHY_ERR_LIB;
    if (rc == CHY_OK) {
      const double *f[6];
      for (int a = 0; a < 6; a++) f[a] = ctx->out_host + a * nc;
      rc = chyo_write_fields(&ctx->prm, st.c_str(), in.c_str(), f);
    }
    ctx->out_rc = rc;
  };
  try {
    ctx->out_thread = std::thread(write);
  } catch (...) {  // no thread to be had: write now, the result waits for chy_output_wait
    write();
  }
  return CHY_OK;
}

int chy_output_wait(chy_ctx *ctx)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  return out_join(ctx);
}

int chy_get_rhs(chy_ctx *ctx, double *b)
{
  if (!ctx || !b) return CHY_ERR_ARG_WRONG;
  CHY_HIP(hipMemcpy(b, ctx->b, sizeof(double) * ctx->g.nc, hipMemcpyDeviceToHost));
  return CHY_OK;
}

int chy_step(chy_ctx *ctx, chy_step_stats *st)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  return chy::ctx_step(ctx, st);
}

}  // extern "C"

int chy::ctx_rhs(chy_ctx *ctx) { return do_rhs(ctx); }
int chy::ctx_solve(chy_ctx *ctx, chy_solve_stats *st) { return do_solve(ctx, st); }

int chy::ctx_step(chy_ctx *ctx, chy_step_stats *st)
{
  const auto t0 = std::chrono::steady_clock::now();
  chy_step_stats local;
  memset(&local, 0, sizeof local);
  CHY_HIP(hipEventRecord(ctx->ev_step, ctx->stream));
  timer_start(ctx, "efield");
  do_efield(ctx);
  timer_stop(ctx, "efield");
  // the push's counters are read with the sort's status, in one synchronisation
  int rc = particles_push(ctx, true, ctx->prm.mn > 0.0, &local.collisions, true);
  ctx->step++;
  if (rc) return rc;
  rc = particles_sort(ctx);
  if (rc) { ctx->push_pending = false; ctx->push_stats = nullptr; return rc; }
  TRY(particles_push_check(ctx));
  rc = particles_deposit(ctx);
  if (rc) return rc;
  // the particle phase ends on the device at this event (no synchronisation)
  CHY_HIP(hipEventRecord(ctx->ev_particles, ctx->stream));
  const auto t1 = std::chrono::steady_clock::now();
  rc = do_rhs(ctx);
  if (rc) return rc;
  // initial phi of the solve (the adjoint's tape records the previous-phi start)
  const int guess = ctx->adj ? CHY_GUESS_PREVIOUS : ctx->prm.ksp_initial_guess;
  if (guess == CHY_GUESS_ZERO) {
    launch_fill<double>(ctx, ctx->g.nc, ctx->phi, 0.0);
  } else if (guess == CHY_GUESS_EXTRAPOLATE && ctx->phi_prev) {
    const size_t bytes = sizeof(double) * ctx->g.nc;
    if (ctx->phi_prev_valid) {  // phi <- 2 phi - phi_prev, phi_prev <- old phi
      CHY_HIP(hipMemcpyAsync(ctx->t64, ctx->phi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
      launch_axpby<double>(ctx, ctx->g.nc, -1.0, ctx->phi_prev, 2.0, ctx->phi);
      CHY_HIP(hipMemcpyAsync(ctx->phi_prev, ctx->t64, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      CHY_HIP(hipMemcpyAsync(ctx->phi_prev, ctx->phi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
      ctx->phi_prev_valid = true;
    }
  }
  rc = do_solve(ctx, &local.solve);
  if (rc) return rc;
  CHY_HIP(hipStreamSynchronize(ctx->stream));
  const auto t2 = std::chrono::steady_clock::now();
  (void)t1;
  // the particle phase: host start of the step to the device's end of the
  // deposit (the device is idle when a step starts: every step ends on a
  // synchronisation); the solve: the rest of the step
  float dev_ms = 0.f;
  CHY_HIP(hipEventElapsedTime(&dev_ms, ctx->ev_step, ctx->ev_particles));
  local.ms_total = std::chrono::duration<double, std::milli>(t2 - t0).count();
  local.ms_particles = std::min<double>(dev_ms, local.ms_total);
  local.solve.ms = local.ms_total - local.ms_particles;
  // global diagnostics of the new state, outside the step's timings
  const int64_t di = ctx->prm.diagnostics_interval;
  if (di > 0 && ctx->step % di == 0) TRY(diagnostics(ctx, &local.diag));
  if (st) *st = local;
  return CHY_OK;
}

extern "C" {

int chy_timing_enable(chy_ctx *ctx, int on)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  ctx->timing = on != 0;
  return CHY_OK;
}

int chy_timing_select(chy_ctx *ctx, const char *names)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  ctx->timing_only.clear();
  if (!names) return CHY_OK;
  std::string cur;
  for (const char *c = names;; c++) {
    if (*c == ',' || *c == '\0') {
      if (!cur.empty()) ctx->timing_only.push_back(cur);
      cur.clear();
      if (*c == '\0') break;
    } else if (*c != ' ') {
      cur += *c;
    }
  }
  return CHY_OK;
}

int chy_timing_get(chy_ctx *ctx, const char *name, double *total_ms, int64_t *count)
{
  if (!ctx || !name) return CHY_ERR_ARG_WRONG;
  timers_resolve(ctx);
  auto it = ctx->timers.find(name);
  if (total_ms) *total_ms = it == ctx->timers.end() ? 0.0 : it->second.total_ms;
  if (count) *count = it == ctx->timers.end() ? 0 : it->second.count;
  return CHY_OK;
}

int chy_timing_reset(chy_ctx *ctx)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  timers_resolve(ctx);
  for (auto &kv : ctx->timers) { kv.second.total_ms = 0.0; kv.second.count = 0; }
  return CHY_OK;
}

int chy_copy_kernel(chy_ctx *ctx, int64_t nbytes, int repeat)
{
  if (!ctx || nbytes <= 0 || repeat < 1) return CHY_ERR_ARG_WRONG;
  const int64_t n16 = (nbytes + 15) / 16;
  if ((n16 + kBlock - 1) / kBlock > (int64_t)0x7fffffff) return set_err(ctx, CHY_ERR_ARG_RANGE, "copy too large");
  void *a = nullptr, *b = nullptr;
  CHY_HIP(hipMalloc(&a, n16 * 16));
  if (hipMalloc(&b, n16 * 16) != hipSuccess) { (void)hipFree(a); return set_err(ctx, CHY_ERR_MEM, "copy buffers"); }
  (void)hipMemsetAsync(a, 1, n16 * 16, ctx->stream);
  for (int r = 0; r < repeat; r++) {
    timer_start(ctx, "copy");
    launch_copy(ctx, n16, a, b);
    timer_stop(ctx, "copy");
    timer_start(ctx, "copy1");
    launch_copy1(ctx, n16, a, b);
    timer_stop(ctx, "copy1");
  }
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, hipGetErrorString(e));
  return CHY_OK;
}

}  // extern "C"
