// chy_internal.h -- context layout and the launch wrappers shared between the
// HIP translation units.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <thread>
#include <vector>

#include "chy_device.h"
#include "chybrid.h"

extern "C" int chyo_write_fields(const chy_params *p, const char *stem, const char *insert, const double *const f[6]);

namespace chy {

// Scaled operator coefficients s_ab = r_ab h_ab / detA (reference lhs.c:405-447).
struct SCoef { double sxx, syy, szz, sxy, sxz, syz, syx, szx, szy; };

// Forcing constants (reference rhs.c:199-223).
struct RhsCoef {
  double r[3][3];
  double detA, scale, cth, cG, E0[3];
  double dx, dy, dz;
  // 1/(2 d_a), 1/d_a^2, 1/(4 dx dy), 1/(4 dx dz), 1/(4 dy dz), and (R E0)_a
  double h1[3], h2[3], hm[3], re[3];
};

template <typename T> struct alignas(4 * sizeof(T)) Vec4 { T x, y, z, w; };

// Device buffers of one particle generation (ping-pong pair in the context).
// FP64: one array per field -- x,y,z absolute positions, vx,vy,vz, the
// optional weight, the id and the cell key.
template <typename T>
struct PState {
  T *x = nullptr, *y = nullptr, *z = nullptr;
  T *vx = nullptr, *vy = nullptr, *vz = nullptr;
  T *w = nullptr;          // optional weights
  uint32_t *pid = nullptr;
  int32_t *cell = nullptr; // cell key
};

// Mixed: per ion a 16-byte record (ox, oy, oz, w) of in-cell offsets in
// [0,1) and the weight (1 when unweighted), a 16-byte record (vx, vy, vz, id)
// with the id's bits in the last lane, and the int32 cell key.  An ion moved
// by the sort in slot order is then three reads (two 16-byte records and the
// key) instead of nine 4-byte ones; most of the sort's gathered reads serve a
// few ions each, so their number, not their bytes, sets its speed.
template <>
struct PState<float> {
  Vec4<float> *pos = nullptr, *vel = nullptr;
  int32_t *cell = nullptr;
};

// Per-name kernel timer: event pairs recorded on the context stream and
// resolved lazily (no synchronisation inside the timed work).
struct Timer {
  std::vector<hipEvent_t> a, b;  // event pool
  size_t used = 0;               // pairs recorded since the last resolve
  bool open = false;
  double total_ms = 0.0;
  int64_t count = 0;
};

}  // namespace chy

struct chy_ctx {
  chy_params prm;
  chy::Grid g;
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  bool mixed = false;

  int64_t cap = 0;   // particle capacity
  int64_t np = 0;    // live particles
  bool weighted = false;
  int64_t step = 0;  // collision stream step counter

  // particle state (double or float pairs, one of them used)
  chy::PState<double> pd[2];
  chy::PState<float> pf[2];
  int cur = 0;
  uint32_t *perm = nullptr;     // scatter slot -> old index (unstable order)
  uint32_t *sort_src = nullptr; // when set, the sort records final slot -> old index here
  int32_t *slotkey = nullptr;   // slot -> cell
  uint32_t *counts = nullptr;   // nc + 1
  uint32_t *offsets = nullptr;  // nc + 1
  uint32_t *cursor = nullptr;   // nc
  uint32_t *scan_tmp = nullptr; // block sums for the scan
  uint32_t *maxcount = nullptr; // largest cell occupancy of the last sort
  // neighbour-move sort (the particles are in the stable order of `offsets`
  // and every ion moved at most one cell in each direction)
  bool order_valid = false;     // particle order = stable cell order of `offsets`
  // mixed: move direction of each ion (0..26, 255 beyond a neighbour), written
  // by the moving push from the stable order; dir_valid while it matches the
  // keys (any other change of keys or order clears it)
  uint8_t *dir = nullptr;
  bool dir_valid = false;
  bool keys_moved = false;      // keys changed since the order was made (directions then unknown)
  bool partials_fresh = false;  // the sort made the deposit's cell partials of the current state
  double *phi_prev = nullptr;   // the previous step's phi (-ksp_initial_guess extrapolate)
  bool phi_prev_valid = false;
  uint8_t *trans = nullptr;     // [27][nc] ions of each old cell per move direction
  uint16_t *info = nullptr;     // per ion: direction << 8 | rank among the cell's ions of that direction
  uint16_t *nbase = nullptr;    // [27][nc] start of each (old cell, direction) group in its new cell
  uint32_t *offs_alt = nullptr; // nc + 1: the new offsets while the old ones are read
  uint32_t *sortflag = nullptr; // set when a move is not to a neighbour or a count overflows
  void *partials = nullptr;     // per-cell corner partials [32][nc] (T)
  double *sred = nullptr;       // collision-normaliser block partials
  double *S_dev = nullptr;      // collision normaliser (device scalar)
  unsigned long long *cstats = nullptr; // accepted, refused, capped, nonfinite

  // grid fields
  void *mom[4] = {nullptr, nullptr, nullptr, nullptr}; // n, Gx, Gy, Gz (T)
  void *efield[1] = {nullptr};  // node E, interleaved {Ex, Ey, Ez, 0} per node (T)
  double *phi = nullptr, *b = nullptr, *r64 = nullptr, *t64 = nullptr;
  bool phi_valid = false;
  bool moments_valid = false;

  // Krylov workspace (T = double in FP64 mode, float in mixed mode)
  int64_t restart = 30;
  void *V = nullptr, *Z = nullptr, *wv = nullptr;
  double *red = nullptr;        // reduction partials [nblk][restart+2]
  double *red_out = nullptr;    // reduced values [red_out_n]
  int64_t red_out_n = 0;
  double *rpl = nullptr;        // per-stencil-block partials of the outer residual's |r|^2
  // Coherent pinned host mirror of red_out: every reduction that writes
  // red_out also writes its value here (red_mirror is the device's address of
  // it), so reading a reduced value back is a stream synchronisation, with no
  // device-to-host copy
  double *h_red = nullptr;
  double *red_mirror = nullptr;
  // Coherent pinned status words (the sort's checks, the push's counters),
  // written by a one-block kernel at the device address status_dev, so a
  // read-back costs one stream synchronisation and no copies
  uint64_t *h_status = nullptr;
  uint64_t *status_dev = nullptr;
  // The last word of h_status is a sequence number: each kernel that writes
  // coherent host memory for a read-back (the mirrored combines, the status
  // kernel) stores its launch's number there after its values, so the host
  // waits for exactly that store (wait_seq) instead of draining the stream.
  uint64_t seq_issued = 0;
  hipEvent_t ev_step = nullptr, ev_particles = nullptr;  // a step's start and the end of its particle phase
  double *push_part = nullptr;  // the FP32 push's per-block collision tallies [kPushBlocks][4], then [64][4]
  bool push_pending = false;    // a push's counters not yet read (deferred to the sort's read-back)
  int64_t push_Nc = 0;
  chy_collision_stats *push_stats = nullptr;
  double *h_scale = nullptr;    // pinned: squared scales of the Krylov basis (GMRES)
  // FFT preconditioner
  void *fft_buf = nullptr;      // complex work array
  void *pc_buf = nullptr;       // z-line / Chebyshev / two-level scratch, 3 vectors (T)
  void *pc2_buf = nullptr;      // two-level FFT preconditioner: x-y fields (T)
  double *pc2_h = nullptr;      // its double workspace: column sums, partial sums
  void *zl_buf = nullptr;       // fft2z: z-line factors (4 nc values of T)
  double *zl_col = nullptr;     // fft2z: per-column Sherman-Morrison scalars
  // asynchronous field output: the six fields snapshotted on the device (6 nc
  // f64), copied on out_stream into pinned out_host, written by out_thread
  double *out_dev = nullptr, *out_host = nullptr;
  hipStream_t out_stream = nullptr;
  hipEvent_t out_snap = nullptr, out_copied = nullptr;
  std::thread out_thread;
  int out_rc = 0;
  double *tw[3] = {nullptr, nullptr, nullptr};   // twiddles per axis (complex)
  double *trig[3] = {nullptr, nullptr, nullptr}; // cos/sin of 2 pi k / N per axis
  bool fft_ok = false;

  bool timing = false;
  std::vector<std::string> timing_only;  // when not empty, the only timers recorded
  std::map<std::string, chy::Timer> timers;

  double *diag_part = nullptr;  // diagnostics: block partials, then the sums (diag_out)
  double *diag_out = nullptr;
  void *mg = nullptr;           // multigrid levels (chy_mg.hip), -pc_type mg
  bool solve_transpose = false; // the Krylov solve runs on A(n)^T (adjoint)
  void *adj = nullptr;          // adjoint tape (chy_adjoint.hip)
};

namespace chy {

constexpr int kRedBlocks = 1024;  // fixed reduction grid: deterministic sums
constexpr int kStatusWords = 16;  // chy_ctx::h_status
constexpr int kSeqWord = kStatusWords - 1;  // its sequence word
constexpr int kPushBlocks = 65536;  // the push's largest grid
constexpr int kBlock = 256;

inline int set_err(chy_ctx *c, int code, const std::string &msg)
{
  if (c) c->err = msg;
  return code;
}

#define CHY_HIP(call)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return chy::set_err(ctx, CHY_ERR_LIB, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

// device allocation helpers
template <typename T>
int dalloc(chy_ctx *ctx, T **p, int64_t n)
{
  if (n <= 0) { *p = nullptr; return CHY_OK; }
  const hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess)
    return set_err(ctx, CHY_ERR_MEM, "hipMalloc of " + std::to_string(sizeof(T) * (size_t)n) + " bytes: " + hipGetErrorString(e));
  return CHY_OK;
}


template <typename T>
int alloc_state(chy_ctx *ctx, PState<T> &s, int64_t cap, bool w)
{
  TRY(dalloc(ctx, &s.x, cap)); TRY(dalloc(ctx, &s.y, cap)); TRY(dalloc(ctx, &s.z, cap));
  TRY(dalloc(ctx, &s.vx, cap)); TRY(dalloc(ctx, &s.vy, cap)); TRY(dalloc(ctx, &s.vz, cap));
  if (w) TRY(dalloc(ctx, &s.w, cap));
  TRY(dalloc(ctx, &s.pid, cap));
  TRY(dalloc(ctx, &s.cell, cap));
  return CHY_OK;
}

template <typename T>
void free_state(PState<T> &s)
{
  void *ps[] = {s.x, s.y, s.z, s.vx, s.vy, s.vz, s.w, s.pid, s.cell};
  for (void *p : ps) if (p) (void)hipFree(p);
  s = PState<T>();
}

template <>
inline int alloc_state<float>(chy_ctx *ctx, PState<float> &s, int64_t cap, bool)
{
  TRY(dalloc(ctx, &s.pos, cap)); TRY(dalloc(ctx, &s.vel, cap));
  TRY(dalloc(ctx, &s.cell, cap));
  return CHY_OK;
}

template <>
inline void free_state<float>(PState<float> &s)
{
  void *ps[] = {s.pos, s.vel, s.cell};
  for (void *p : ps) if (p) (void)hipFree(p);
  s = PState<float>();
}

// timers (no-ops unless ctx->timing)
void timer_start(chy_ctx *ctx, const char *name);
// Waits until the device has stored sequence number s (or a later one) in the
// context's coherent sequence word.
int wait_seq(chy_ctx *ctx, uint64_t s);
// The device address of the sequence word, and the next launch's number.
inline unsigned long long *seq_word(chy_ctx *ctx) { return (unsigned long long *)(ctx->status_dev + kSeqWord); }
inline unsigned long long next_seq(chy_ctx *ctx) { return (unsigned long long)++ctx->seq_issued; }
void timer_stop(chy_ctx *ctx, const char *name);

// ---- grid kernels (chy_grid.hip) ----
SCoef full_scoef(const double d[3], const double K[3]);
void uniform_row(int lhs_type, const double d[3], const double K[3], double nbar, double v[19]);
template <typename TN, typename TX, typename TY>
void launch_lhs_apply(chy_ctx *ctx, int lhs_type, bool transpose, const TN *n, const TX *x, TY *y);
template <typename TN, typename TX>
void launch_residual(chy_ctx *ctx, int lhs_type, const TN *n, const double *b, const TX *x, double *r);
template <typename TN, typename TY>
void launch_residual_norm(chy_ctx *ctx, int lhs_type, const TN *n, const double *b, const double *x, TY *r,
                          double *pl);
template <typename TS>
void launch_add_scaled(chy_ctx *ctx, int64_t n, const TS *x, double a, double *y);
template <typename T>
int64_t launch_residual_sq(chy_ctx *ctx, int lhs_type, const T *n, const T *b, const T *x, T *r, double *pl);
// part[zc * nx * ny + col] = sum over the planes of z chunk zc of (b - A x) at
// the column; pl[blk] = the sum over each stencil block (lhs_zsum_blocks of them);
// when r is given, b - A x itself is also stored there
template <typename T>
void launch_lhs_zsum(chy_ctx *ctx, int lhs_type, const T *n, const T *x, const T *b, double *part, double *pl,
                     T *r = nullptr);
int lhs_zchunks(const Grid &g);
int64_t lhs_zsum_blocks(const Grid &g, bool fp32);  // FP32 or FP64 stencil's grid
template <typename T>
void launch_efield(chy_ctx *ctx, const double *phi, T *E4);
template <typename T>
void launch_rhs(chy_ctx *ctx, int rhs_type, const T *n, const T *gx, const T *gy, const T *gz, double *b);
// reductions: out[k] = sum_i a_k[i] * b[i] for k < nv (a_k = A + k*stride)
template <typename T>
void launch_mdot(chy_ctx *ctx, int nv, const T *A, int64_t stride, const T *b, int64_t n, double *out_dev,
                 const double *scale = nullptr);  // out[k] *= scale[k] when given (a host array, passed by value)
template <typename T>
void launch_sum(chy_ctx *ctx, const T *a, int64_t n, double *out_dev);
// y -= sum_k h_k V_k (h on device, FP64) ; optional normalisation
template <typename T>
void launch_mgs_update(chy_ctx *ctx, int nv, const T *V, int64_t stride, const double *h, T *w, int64_t n);
template <typename T>
void launch_mgs_update_host(chy_ctx *ctx, int nv, const T *V, int64_t stride, const double *h_host, T *w, int64_t n);
template <typename T>
void launch_mgs_norm(chy_ctx *ctx, int nv, const T *V, int64_t stride, const double *h, T *w, int64_t n,
                     double *norm2_dev);
// Gram-Schmidt passes (vector loads; fall back to the kernels above for
// nv > 8 or n not a multiple of 4): dots V^T w; w -= V h then V^T w;
// w -= V h then ||w||^2
template <typename T>
void launch_gs_dots(chy_ctx *ctx, int nv, const T *V, int64_t stride, const T *w, int64_t n, double *out_dev,
                    const double *scale = nullptr);  // host scales, as launch_mdot
template <typename T>
void launch_gs_update_dots(chy_ctx *ctx, int nv, const T *V, int64_t stride, const double *h, T *w, int64_t n,
                           double *out_dev, const double *scale = nullptr);
template <typename T>
void launch_gs_update_norm(chy_ctx *ctx, int nv, const T *V, int64_t stride, const double *h, T *w, int64_t n,
                           double *norm2_dev);
template <typename T>
void launch_axpby(chy_ctx *ctx, int64_t n, double a, const T *x, double b, T *y);
template <typename TS, typename TD>
void launch_convert(chy_ctx *ctx, int64_t n, const TS *src, double scale, TD *dst);
template <typename T>
void launch_shift(chy_ctx *ctx, int64_t n, T *v, const double *shift_dev, double scale);
template <typename T>
void launch_fill(chy_ctx *ctx, int64_t n, T *v, T val);
template <typename T>
void launch_jacobi(chy_ctx *ctx, int lhs_type, const T *n, const T *r, T *z);
// coefficients of the z-averaged operator, C[9][nx*ny] (two-level FFT preconditioner)
template <typename T>
void launch_zcoef(chy_ctx *ctx, int lhs_type, const T *n, T *C);
// z-line block Jacobi (scratch cp, q: nc each) and Chebyshev-Jacobi (scratch
// res, dv, Ad: nc each) preconditioners
template <typename T>
void launch_zline(chy_ctx *ctx, int lhs_type, const T *n, const T *r, T *z, T *cp, T *q);
template <typename T>
void launch_cheby(chy_ctx *ctx, int lhs_type, bool transpose, int degree, const T *n, const T *r, T *z, T *res,
                  T *dv, T *Ad);
void launch_combine_sum(chy_ctx *ctx, const double *partials, int nblk, int nv, double *out);
void launch_copy(chy_ctx *ctx, int64_t n16, const void *a, void *b);   // grid-stride, 16384 blocks
void launch_copy1(chy_ctx *ctx, int64_t n16, const void *a, void *b);  // one element per thread
RhsCoef rhs_coef(const chy_ctx *ctx);
// adjoint pieces (FP64): out += alpha d(mu^T A(n) phi)/dn; transposed forcing;
// central gradient of n into g3[3][nc]; transposed E-field stencil
void launch_lhs_dn(chy_ctx *ctx, const double *mu, const double *phi, double alpha, double *out);
void launch_rhs_full_adj(chy_ctx *ctx, const double *mu, double *ln, double *lgx, double *lgy, double *lgz);
void launch_grad3(chy_ctx *ctx, const double *n, double *g3);
void launch_efield_adj(chy_ctx *ctx, const double *lex, const double *ley, const double *lez, double *lphi);

// ---- particle kernels (chy_particles.hip) ----
int particles_upload(chy_ctx *ctx, int64_t np, const double *pos, const double *vel, const double *w,
                     const uint32_t *pid);
int particles_download(chy_ctx *ctx, double *pos, double *vel, double *w, uint32_t *pid);
int particles_init(chy_ctx *ctx);
// defer: leave the counters to the next status read-back (the sort's), which
// checks them first; particles_push_check reads them if still pending
int particles_push(chy_ctx *ctx, bool move, bool collide, chy_collision_stats *st, bool defer = false);
int particles_push_check(chy_ctx *ctx);
int particles_move_only(chy_ctx *ctx);
int particles_sort(chy_ctx *ctx);
int particles_deposit(chy_ctx *ctx);
int particles_normaliser(chy_ctx *ctx);

// deposit of an arbitrary cell-sorted FP64 state (segments `offs`) into
// out[4] = {sum w W, sum w W vx, sum w W vy, sum w W vz}
void deposit_state(chy_ctx *ctx, const PState<double> &st, const uint32_t *offs, bool weighted, double *const out[4]);

// ---- diagnostics (chy_diag.hip) ----
int diagnostics(chy_ctx *ctx, chy_diag_stats *d);
int diag_scratch_init(chy_ctx *ctx);  // this device's static scratch address (context creation)
int64_t diag_elems();  // device doubles for diag_part + diag_out
int diag_parts();      // of which block partials (diag_out follows them)

// ---- multigrid preconditioner (chy_mg.hip) ----
int mg_setup(chy_ctx *ctx);    // levels and their buffers (context creation)
void mg_release(chy_ctx *ctx);
int mg_levels(const chy_ctx *ctx);
template <typename T>
int mg_prepare(chy_ctx *ctx, const T *n);  // per solve: level densities, z-line factors
template <typename T>
int mg_apply(chy_ctx *ctx, const T *r, T *z);  // z = one V-cycle on r

// ---- context-level operations (chy_ctx.hip) ----
int ctx_step(chy_ctx *ctx, chy_step_stats *st);  // one step of the loop body
int ctx_rhs(chy_ctx *ctx);                       // forcing from the current moments
int ctx_solve(chy_ctx *ctx, chy_solve_stats *st); // A(n) phi = b (A^T when solve_transpose)
void adjoint_release(chy_ctx *ctx);              // frees the adjoint tape

// ---- FFT preconditioner (chy_fft.hip) ----
int fft_setup(chy_ctx *ctx);
template <typename T>
int fft_pc_apply(chy_ctx *ctx, const double v19[19], const T *r, T *z);
// two-level: fft, then kFft2Its steps of flexible GMRES on the z-averaged
// residual with the z-averaged operator (coefficients from launch_zcoef in
// pc2_buf); w3 is 3-D scratch
constexpr int kFft2Its = 2;
int64_t fft2_buf_elems(const chy_ctx *ctx);  // elements of T in pc2_buf
int64_t fft2_h_elems(const chy_ctx *ctx);    // doubles in pc2_h
// fft2z: the z-line correction of the fft2 residual.  zbuf: 4 nc values of T
// (factors), colbuf: 2 (nx ny) doubles, made by launch_zline_factor once per
// solve; launch_zline_apply adds the z-line solve of res to z (res is used as
// scratch); launch_residual_t: res = b - A(n) x in precision T.
template <typename T>
void launch_zline_factor(chy_ctx *ctx, int lhs_type, const T *n, T *zbuf, double *colbuf);
// the full operator's z-averaged coefficients (Cz, as launch_zcoef) and/or
// z-line factors (zbuf, colbuf, as launch_zline_factor) in one column walk;
// a null Cz or zbuf skips that part
template <typename T>
void launch_zcolumn(chy_ctx *ctx, const T *n, T *Cz, T *zbuf, double *colbuf);
template <typename T>
void launch_zline_apply(chy_ctx *ctx, const T *zbuf, const double *colbuf, T *res, T *z);
// fft2z's z-line stage with its 2-D correction folded in, where it applies
// (the full operator with B along z): z += e2 + zline(res - A(n) e2), e2 =
// sum_q y2[q] Z2[q] constant along z, res = r - A(n) z on entry
template <typename T>
void launch_zline_apply_e2(chy_ctx *ctx, const T *zbuf, const double *colbuf, T *res, T *z, const T *n, const T *Z2,
                           const double *y2);
bool zline_e2_applies(chy_ctx *ctx);
template <typename T>
void launch_residual_t(chy_ctx *ctx, int lhs_type, const T *n, const T *b, const T *x, T *res);
// the same pieces on an explicit grid g with spacings d (multigrid levels):
// z-line factors, z += scale zline^-1 res, res = b - A(n) x
template <typename T>
void launch_zline_factor_g(chy_ctx *ctx, const Grid &g, const double d[3], int lhs_type, const T *n, T *zbuf,
                           double *colbuf);
template <typename T>
void launch_zline_apply_g(chy_ctx *ctx, const Grid &g, const T *zbuf, const double *colbuf, T *res, T *z,
                          double scale);
// the same, one wave per column on grids of few columns (res is only read)
template <typename T>
void launch_zline_apply_coarse(chy_ctx *ctx, const Grid &g, const T *zbuf, const double *colbuf, T *res, T *z,
                               double scale);
template <typename T>
void launch_residual_g(chy_ctx *ctx, const Grid &g, const double d[3], int lhs_type, const T *n, const T *b,
                       const T *x, T *res);
// The two-level preconditioner z = fft2(r).  With zline set it goes on to
// fft2z's z-line stage, z += zline(r - A(n) z), in the fused form where that
// applies, and says so in *zline_done (else the caller runs the stage); w3 is
// a scratch vector for the stage's residual.
template <typename T>
int fft2_pc_apply(chy_ctx *ctx, const double v19[19], const T *r, T *z, T *w3, bool zline = false,
                  bool *zline_done = nullptr);

}  // namespace chy
