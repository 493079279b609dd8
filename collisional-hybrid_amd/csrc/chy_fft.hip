// chy_fft.hip -- FFT preconditioner for the periodic 19-point operator.
//
// P = A(nbar), the operator of the mean density, is a constant-coefficient
// periodic stencil with a_o = a_{-o}, hence diagonalised by the 3-D DFT with
// real symbol lam(k) = sum_o a_o cos(k.o).  z = P^+ r is computed as
// IFFT(FFT(r) / lam) with the k = 0 mode removed (the operator's constant
// null space).  All passes are LDS-resident batched radix-2 transforms:
// forward passes use decimation in frequency (natural in, bit-reversed out)
// and inverse passes decimation in time (bit-reversed in, natural out), so
// no permutation pass is needed; the z pass fuses forward FFT, the symbol
// division and the inverse FFT.  Five HBM passes: x-fwd (real in), y-fwd,
// z-fwd/scale/z-inv, y-inv, x-inv (real out).  Grid sizes: powers of two.
#include <cmath>

#include "chy_internal.h"

namespace chy {

template <typename T> struct cpx { T x, y; };

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

__device__ __forceinline__ int bitrev(int v, int logn) { return (int)(__brev((unsigned)v) >> (32 - logn)); }

constexpr int kFftMaxN = 2048;  // longest axis handled by one LDS-resident line

// Lines per block: the largest power of two whose padded LDS footprint
// B*(n+1) complex values stays within 64 KB, and at most `avail`.
template <typename T>
static int fft_batch(int n, int64_t avail)
{
  const int64_t E = 65536 / (int64_t)(2 * sizeof(T));
  int B = 1;
  while ((int64_t)(2 * B) * (n + 1) <= E && 2 * B <= avail) B *= 2;
  return B;
}

// One batch of lines in LDS: s[b*(n+1) + t].
template <typename T>
__device__ __forceinline__ void dif_fwd(cpx<T> *s, int B, int n, int logn, const double *tw)
{
  // tw[2j], tw[2j+1] = cos, sin of -2 pi j / n  (j < n/2)
  for (int len = n; len >= 2; len >>= 1) {
    const int half = len >> 1, step = n / len;
    for (int m = threadIdx.x; m < B * (n >> 1); m += blockDim.x) {
      const int b = m / (n >> 1), r = m % (n >> 1);
      const int grp = r / half, k = r % half;
      const int i0 = b * (n + 1) + grp * len + k, i1 = i0 + half;
      const cpx<T> u = s[i0], v = s[i1];
      const cpx<T> w = {(T)tw[2 * (k * step)], (T)tw[2 * (k * step) + 1]};
      s[i0] = {u.x + v.x, u.y + v.y};
      s[i1] = cmul<T>({u.x - v.x, u.y - v.y}, w);
    }
    __syncthreads();
  }
}

template <typename T>
__device__ __forceinline__ void dit_inv(cpx<T> *s, int B, int n, int logn, const double *tw)
{
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1, step = n / len;
    for (int m = threadIdx.x; m < B * (n >> 1); m += blockDim.x) {
      const int b = m / (n >> 1), r = m % (n >> 1);
      const int grp = r / half, k = r % half;
      const int i0 = b * (n + 1) + grp * len + k, i1 = i0 + half;
      const cpx<T> w = {(T)tw[2 * (k * step)], -(T)tw[2 * (k * step) + 1]};  // conjugate: +2 pi i
      const cpx<T> u = s[i0], v = cmul<T>(s[i1], w);
      s[i0] = {u.x + v.x, u.y + v.y};
      s[i1] = {u.x - v.x, u.y - v.y};
    }
    __syncthreads();
  }
}

// Pass along x (contiguous lines).  MODE 0: real in -> forward -> complex out.
// MODE 1: complex in -> inverse -> real out (real part times scale).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_fft_x(int n, int logn, int nlines, int B, const T *__restrict__ rin,
                                               cpx<T> *__restrict__ buf, T *__restrict__ rout, const double *__restrict__ tw,
                                               double scale)
{
  extern __shared__ unsigned char smem_raw[];
  cpx<T> *s = reinterpret_cast<cpx<T> *>(smem_raw);
  const int64_t l0 = (int64_t)blockIdx.x * B;
  const int nb = (int)((nlines - l0) < B ? (nlines - l0) : B);
  for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
    const int b = e / n, t = e % n;
    const int64_t a = (l0 + b) * n + t;
    if (MODE == 0) s[b * (n + 1) + t] = {rin[a], (T)0};
    else s[b * (n + 1) + t] = buf[a];
  }
  __syncthreads();
  if (MODE == 0) dif_fwd<T>(s, nb, n, logn, tw);
  else dit_inv<T>(s, nb, n, logn, tw);
  for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
    const int b = e / n, t = e % n;
    const int64_t a = (l0 + b) * n + t;
    if (MODE == 0) buf[a] = s[b * (n + 1) + t];
    else rout[a] = (T)((double)s[b * (n + 1) + t].x * scale);
  }
}

// Pass along a strided axis: line (a, c) at base a + c*cstride, element t at
// + t*estride; a block holds B consecutive a.  MODE 0 forward, 1 inverse,
// 2 forward + symbol division + inverse (z axis).
struct Symbol {
  double v[19];
  int off[19][3];
  int nx, ny, nz, lx, ly, lz;  // sizes and log2 sizes
  double inv_scale;            // 1 / (nx ny nz)
};

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_fft_s(int n, int logn, int64_t A, int64_t estride, int64_t cstride, int B,
                                               cpx<T> *__restrict__ buf, const double *__restrict__ tw, Symbol sym,
                                               const double *__restrict__ trx, const double *__restrict__ try_,
                                               const double *__restrict__ trz)
{
  extern __shared__ unsigned char smem_raw[];
  cpx<T> *s = reinterpret_cast<cpx<T> *>(smem_raw);
  const int64_t blocks_per_c = A / B;
  const int64_t c = blockIdx.x / blocks_per_c;
  const int64_t a0 = (blockIdx.x % blocks_per_c) * B;
  const int64_t base = a0 + c * cstride;
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    const int t = e / B, b = e % B;
    s[b * (n + 1) + t] = buf[base + b + (int64_t)t * estride];
  }
  __syncthreads();
  if (MODE == 1) {
    dit_inv<T>(s, B, n, logn, tw);
  } else {
    dif_fwd<T>(s, B, n, logn, tw);
    if (MODE == 2) {
      // line a = ix + nx*iy (positions hold bit-reversed x and y frequencies),
      // element t holds bit-reversed z frequency
      for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
        const int t = e / B, b = e % B;
        const int64_t a = a0 + b;
        const int kx = bitrev((int)(a % sym.nx), sym.lx);
        const int ky = bitrev((int)(a / sym.nx), sym.ly);
        const int kz = bitrev(t, sym.lz);
        const double cs[3] = {trx[2 * kx], try_[2 * ky], trz[2 * kz]};
        const double sn[3] = {trx[2 * kx + 1], try_[2 * ky + 1], trz[2 * kz + 1]};
        double lam = 0.0;
#pragma unroll
        for (int q = 0; q < 19; q++) {
          double C = 1.0, S = 0.0;
#pragma unroll
          for (int ax = 0; ax < 3; ax++) {
            const int o = sym.off[q][ax];
            if (o != 0) {
              const double ss = o * sn[ax];
              const double nC = C * cs[ax] - S * ss, nS = S * cs[ax] + C * ss;
              C = nC; S = nS;
            }
          }
          lam += sym.v[q] * C;
        }
        cpx<T> &z = s[b * (n + 1) + t];
        if (lam == 0.0 || (kx == 0 && ky == 0 && kz == 0)) {
          z = {(T)0, (T)0};
        } else {
          const double f = sym.inv_scale / lam;
          z = {(T)((double)z.x * f), (T)((double)z.y * f)};
        }
      }
      __syncthreads();
      dit_inv<T>(s, B, n, logn, tw);
    }
  }
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    const int t = e / B, b = e % B;
    buf[base + b + (int64_t)t * estride] = s[b * (n + 1) + t];
  }
}

static bool pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
static int ilog2(int64_t n) { int l = 0; while ((1ll << l) < n) l++; return l; }

int fft_setup(chy_ctx *ctx)
{
  const int64_t N[3] = {ctx->g.nx, ctx->g.ny, ctx->g.nz};
  ctx->fft_ok = pow2(N[0]) && pow2(N[1]) && pow2(N[2]) && N[0] <= kFftMaxN && N[1] <= kFftMaxN && N[2] <= kFftMaxN;
  if (!ctx->fft_ok) return CHY_OK;
  for (int a = 0; a < 3; a++) {
    std::vector<double> tw(N[a]), tr(2 * N[a]);
    for (int64_t j = 0; j < N[a] / 2; j++) {
      tw[2 * j] = cos(-2.0 * M_PI * (double)j / (double)N[a]);
      tw[2 * j + 1] = sin(-2.0 * M_PI * (double)j / (double)N[a]);
    }
    for (int64_t k = 0; k < N[a]; k++) {
      tr[2 * k] = cos(2.0 * M_PI * (double)k / (double)N[a]);
      tr[2 * k + 1] = sin(2.0 * M_PI * (double)k / (double)N[a]);
    }
    CHY_HIP(hipMalloc(&ctx->tw[a], sizeof(double) * std::max<int64_t>(2, N[a])));
    CHY_HIP(hipMalloc(&ctx->trig[a], sizeof(double) * 2 * N[a]));
    CHY_HIP(hipMemcpy(ctx->tw[a], tw.data(), sizeof(double) * N[a], hipMemcpyHostToDevice));
    CHY_HIP(hipMemcpy(ctx->trig[a], tr.data(), sizeof(double) * 2 * N[a], hipMemcpyHostToDevice));
  }
  return CHY_OK;
}

template <typename T>
int fft_pc_apply(chy_ctx *ctx, const double v19[19], const T *r, T *z)
{
  const int nx = ctx->g.nx, ny = ctx->g.ny, nz = ctx->g.nz;
  cpx<T> *buf = (cpx<T> *)ctx->fft_buf;
  Symbol sym;
  static const int off[19][3] = {
    {+1, 0, 0}, {-1, 0, 0}, {0, +1, 0}, {0, -1, 0}, {0, 0, +1}, {0, 0, -1},
    {+1, +1, 0}, {+1, -1, 0}, {-1, +1, 0}, {-1, -1, 0},
    {+1, 0, +1}, {+1, 0, -1}, {-1, 0, +1}, {-1, 0, -1},
    {0, +1, +1}, {0, +1, -1}, {0, -1, +1}, {0, -1, -1},
    {0, 0, 0}};
  for (int q = 0; q < 19; q++) {
    sym.v[q] = v19[q];
    for (int a = 0; a < 3; a++) sym.off[q][a] = off[q][a];
  }
  sym.nx = nx; sym.ny = ny; sym.nz = nz;
  sym.lx = ilog2(nx); sym.ly = ilog2(ny); sym.lz = ilog2(nz);
  sym.inv_scale = 1.0 / ((double)nx * ny * nz);
  // x forward (real in)
  {
    const int n = nx, B = fft_batch<T>(n, (int64_t)ny * nz);
    const int nlines = ny * nz;
    const unsigned nb = (unsigned)((nlines + B - 1) / B);
    const size_t sh = (size_t)B * (n + 1) * sizeof(cpx<T>);
    hipLaunchKernelGGL((k_fft_x<T, 0>), nb, kBlock, sh, ctx->stream, n, ilog2(n), nlines, B, r, buf, (T *)nullptr,
                       ctx->tw[0], 1.0);
  }
  // y forward: lines (ix, iz); A = nx contiguous, c = iz with stride nx*ny
  {
    const int n = ny, B = fft_batch<T>(n, nx);
    const unsigned nb = (unsigned)((nx / B) * nz);
    const size_t sh = (size_t)B * (n + 1) * sizeof(cpx<T>);
    hipLaunchKernelGGL((k_fft_s<T, 0>), nb, kBlock, sh, ctx->stream, n, ilog2(n), (int64_t)nx, (int64_t)nx,
                       (int64_t)nx * ny, B, buf, ctx->tw[1], sym, ctx->trig[0], ctx->trig[1], ctx->trig[2]);
  }
  // z forward + symbol + inverse: lines a = ix + nx*iy, estride nx*ny
  {
    const int n = nz;
    const int64_t A = (int64_t)nx * ny;
    const int B = fft_batch<T>(n, A);
    const unsigned nb = (unsigned)(A / B);
    const size_t sh = (size_t)B * (n + 1) * sizeof(cpx<T>);
    hipLaunchKernelGGL((k_fft_s<T, 2>), nb, kBlock, sh, ctx->stream, n, ilog2(n), A, A, (int64_t)0, B, buf,
                       ctx->tw[2], sym, ctx->trig[0], ctx->trig[1], ctx->trig[2]);
  }
  // y inverse
  {
    const int n = ny, B = fft_batch<T>(n, nx);
    const unsigned nb = (unsigned)((nx / B) * nz);
    const size_t sh = (size_t)B * (n + 1) * sizeof(cpx<T>);
    hipLaunchKernelGGL((k_fft_s<T, 1>), nb, kBlock, sh, ctx->stream, n, ilog2(n), (int64_t)nx, (int64_t)nx,
                       (int64_t)nx * ny, B, buf, ctx->tw[1], sym, ctx->trig[0], ctx->trig[1], ctx->trig[2]);
  }
  // x inverse (real out)
  {
    const int n = nx, B = fft_batch<T>(n, (int64_t)ny * nz);
    const int nlines = ny * nz;
    const unsigned nb = (unsigned)((nlines + B - 1) / B);
    const size_t sh = (size_t)B * (n + 1) * sizeof(cpx<T>);
    hipLaunchKernelGGL((k_fft_x<T, 1>), nb, kBlock, sh, ctx->stream, n, ilog2(n), nlines, B, (const T *)nullptr, buf,
                       z, ctx->tw[0], 1.0);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, std::string("fft: ") + hipGetErrorString(e));
  return CHY_OK;
}

template int fft_pc_apply<double>(chy_ctx *, const double *, const double *, double *);
template int fft_pc_apply<float>(chy_ctx *, const double *, const float *, float *);

}  // namespace chy
