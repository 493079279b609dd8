// chy_diag.hip -- global diagnostics of the state, summed in FP64 on the
// device over fixed grids (deterministic): ion count and total weight, ion
// momentum, ion kinetic energy, and the energy of the electrostatic field of
// the potential.  The reference prints per-step collision counts
// (particles.c:525,620) and no global conservation figures; these are the
// energy and momentum diagnostics of the GPU path.
//
//   momentum        P = m_i sum_p w_p v_p                          [kg m/s]
//   kinetic energy  K = 1/2 m_i sum_p w_p |v_p|^2                  [J]
//   field energy    W = 1/2 eps0 sum_nodes |grad phi|^2 dx dy dz   [J]
//                   (central differences, periodic, as the node field
//                   E = E0 - grad phi that the push gathers; the imposed
//                   uniform E0 adds the constant 1/2 eps0 |E0|^2 L^3)
// with w_p the ion's weight (1 when unweighted): one simulation ion per unit
// weight, as the density moments count them.
#include "chy_internal.h"

namespace chy {

namespace {

constexpr int kDiagBlocks = 1024;  // fixed reduction grid
constexpr int kIonTerms = 5;       // sum w, sum w v (3), sum w |v|^2
constexpr double kEps0 = 8.8541878128e-12;  // vacuum permittivity [F/m], CODATA 2018

__device__ __forceinline__ double dwave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block sums of NV values in a fixed order, written to part[blockIdx.x * NV + k].
template <int NV>
__device__ __forceinline__ void block_store(double (&a)[NV], double *__restrict__ part)
{
  __shared__ double sh[NV][4];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = dwave_sum(a[k]);
    if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    part[(int64_t)blockIdx.x * NV + k] = (sh[k][0] + sh[k][1]) + (sh[k][2] + sh[k][3]);
  }
}

__global__ __launch_bounds__(256) void k_diag_ions_fp64(int64_t np, PState<double> st, bool weighted,
                                                        double *__restrict__ part)
{
  double a[kIonTerms] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
    const double w = weighted ? st.w[p] : 1.0, vx = st.vx[p], vy = st.vy[p], vz = st.vz[p];
    a[0] += w;
    a[1] += w * vx; a[2] += w * vy; a[3] += w * vz;
    a[4] += w * ((vx * vx + vy * vy) + vz * vz);
  }
  block_store<kIonTerms>(a, part);
}

// Mixed state: the weight is the fourth lane of the position record, the
// velocity the first three of the velocity record; products formed in FP64.
__global__ __launch_bounds__(256) void k_diag_ions_mixed(int64_t np, PState<float> st, double *__restrict__ part)
{
  double a[kIonTerms] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
    const Vec4<float> v = st.vel[p];
    const double w = (double)st.pos[p].w, vx = (double)v.x, vy = (double)v.y, vz = (double)v.z;
    a[0] += w;
    a[1] += w * vx; a[2] += w * vy; a[3] += w * vz;
    a[4] += w * ((vx * vx + vy * vy) + vz * vz);
  }
  block_store<kIonTerms>(a, part);
}

// sum over nodes of |grad phi|^2 with periodic central differences.
__global__ __launch_bounds__(256) void k_diag_field(Grid g, double hx, double hy, double hz,
                                                    const double *__restrict__ phi, double *__restrict__ part)
{
  double a[1] = {0.0};
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((int64_t)g.nx * g.ny));
    const double ex = (phi[lin(g, i + 1, j, k)] - phi[lin(g, i - 1, j, k)]) * hx;
    const double ey = (phi[lin(g, i, j + 1, k)] - phi[lin(g, i, j - 1, k)]) * hy;
    const double ez = (phi[lin(g, i, j, k + 1)] - phi[lin(g, i, j, k - 1)]) * hz;
    a[0] += (ex * ex + ey * ey) + ez * ez;
  }
  block_store<1>(a, part);
}

// out[0..5] = {weight, momentum (3), kinetic energy, field energy} from the
// six sums, with the scalings the host applies in diagnostics().
__global__ void k_diag_scale(const double *__restrict__ sums, double mi, double fscale, double *__restrict__ out)
{
  if (threadIdx.x != 0) return;
  out[0] = sums[0];
  for (int a = 0; a < 3; a++) out[1 + a] = mi * sums[1 + a];
  out[4] = (0.5 * mi) * sums[4];
  out[5] = fscale * sums[5];
}

// The allocation-free entry point's scratch: the block partials of the six
// sums and the sums themselves (kDiagBlocks * 6 + 6 doubles, 48 KiB), one
// copy per GPU for the life of the process.  Host code uses the per-device
// address from hipGetSymbolAddress, never the symbol itself.
constexpr int kScratchDoubles = kDiagBlocks * (kIonTerms + 1) + kIonTerms + 1;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *scratch_ptr[kMaxDevices];

// Enqueues the ion and field sums of the context's state: block partials in
// part (kDiagBlocks * 6), the six sums in sums.
void diag_sums(chy_ctx *ctx, double *part, double *sums)
{
  const chy_params &p = ctx->prm;
  const int64_t np = ctx->np;
  const unsigned nbi = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kDiagBlocks, (np + kBlock - 1) / kBlock));
  const unsigned nbf =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>(kDiagBlocks, (ctx->g.nc + kBlock - 1) / kBlock));
  if (ctx->mixed)
    hipLaunchKernelGGL(k_diag_ions_mixed, nbi, kBlock, 0, ctx->stream, np, ctx->pf[ctx->cur], part);
  else
    hipLaunchKernelGGL(k_diag_ions_fp64, nbi, kBlock, 0, ctx->stream, np, ctx->pd[ctx->cur], ctx->weighted, part);
  launch_combine_sum(ctx, part, (int)nbi, kIonTerms, sums);
  double *fpart = part + (int64_t)kDiagBlocks * kIonTerms;
  if (ctx->phi_valid) {
    hipLaunchKernelGGL(k_diag_field, nbf, kBlock, 0, ctx->stream, ctx->g, 0.5 / p.d[0], 0.5 / p.d[1], 0.5 / p.d[2],
                       ctx->phi, fpart);
    launch_combine_sum(ctx, fpart, (int)nbf, 1, sums + kIonTerms);
  } else {
    launch_fill<double>(ctx, 1, sums + kIonTerms, 0.0);
  }
}

double field_scale(const chy_params &p) { return (0.5 * kEps0) * (p.d[0] * p.d[1] * p.d[2]); }

}  // namespace

// Address of this device's scratch array (once per device, that device
// current: the context's creation).
int diag_scratch_init(chy_ctx *ctx)
{
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_err(ctx, CHY_ERR_ARG_RANGE, "device index beyond the scratch table");
  if (scratch_ptr[ctx->device]) return CHY_OK;
  double *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_err(ctx, CHY_ERR_LIB, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  scratch_ptr[ctx->device] = ptr;
  return CHY_OK;
}

// Fills d from the current ions and potential (the potential term is zero
// before the first solve).  One host read-back.
int diagnostics(chy_ctx *ctx, chy_diag_stats *d)
{
  const chy_params &p = ctx->prm;
  double *part = ctx->diag_part, *out = ctx->diag_out;
  const int64_t np = ctx->np;
  diag_sums(ctx, part, out);
  double h[kIonTerms + 1];
  CHY_HIP(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  CHY_HIP(hipStreamSynchronize(ctx->stream));
  d->ions = np;
  d->weight = h[0];
  for (int a = 0; a < 3; a++) d->momentum[a] = p.mi * h[1 + a];
  d->kinetic_energy = (0.5 * p.mi) * h[4];
  d->field_energy = field_scale(p) * h[5];
  d->step = ctx->step;
  d->valid = 1;
  return CHY_OK;
}

// The same figures into caller device memory, with the partials in this
// device's static scratch: no allocation, no host synchronisation.
int diagnostics_noalloc(chy_ctx *ctx, double *d_out)
{
  double *base = scratch_ptr[ctx->device];
  if (!base) return set_err(ctx, CHY_ERR_LIB, "the scratch array of this device has no address");
  double *part = base, *sums = base + (int64_t)kDiagBlocks * (kIonTerms + 1);
  diag_sums(ctx, part, sums);
  hipLaunchKernelGGL(k_diag_scale, 1, 64, 0, ctx->stream, (const double *)sums, ctx->prm.mi, field_scale(ctx->prm),
                     d_out);
  CHY_HIP(hipGetLastError());
  return CHY_OK;
}

// Device doubles the diagnostics need: block partials and the sums.
int64_t diag_elems() { return (int64_t)kDiagBlocks * (kIonTerms + 1) + kIonTerms + 1; }
int diag_parts() { return kDiagBlocks * (kIonTerms + 1); }

}  // namespace chy

using namespace chy;

extern "C" int chy_diagnostics(chy_ctx *ctx, chy_diag_stats *d)
{
  if (!ctx || !d) return CHY_ERR_ARG_WRONG;
  return diagnostics(ctx, d);
}

extern "C" int chy_diagnostics_noalloc(chy_ctx *ctx, double *d_out)
{
  if (!ctx) return CHY_ERR_ARG_WRONG;
  if (!d_out) return set_err(ctx, CHY_ERR_ARG_WRONG, "chy_diagnostics_noalloc: no device output array");
  return diagnostics_noalloc(ctx, d_out);
}
