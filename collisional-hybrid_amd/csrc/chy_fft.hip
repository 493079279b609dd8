// chy_fft.hip -- FFT preconditioner for the periodic 19-point operator.
//
// P = A(nbar), the operator of the mean density, is a constant-coefficient
// periodic stencil with a_o = a_{-o}, hence diagonalised by the 3-D DFT with
// real symbol lam(k) = sum_o a_o cos(k.o).  z = P^+ r is computed as
// IFFT(FFT(r) / lam) with the k = 0 mode removed (the operator's constant
// null space).  All passes are LDS-resident batched radix-4 transforms:
// forward passes use decimation in frequency (natural in, bit-reversed out)
// and inverse passes decimation in time (bit-reversed in, natural out), so
// no permutation pass is needed; the z pass fuses forward FFT, the symbol
// division and the inverse FFT.  The x pass is real-to-complex (half
// spectrum, nx/2 complex values per line), so every pass moves as many bytes
// as the real field: x-fwd, y-fwd, z-fwd/scale/z-inv, y-inv, x-inv.  Grid
// sizes: powers of two.
#include <algorithm>
#include <cmath>

#include "chy_internal.h"

namespace chy {

template <typename T> struct cpx { T x, y; };

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

__device__ __forceinline__ int bitrev(int v, int logn) { return (int)(__brev((unsigned)v) >> (32 - logn)); }

constexpr int kFftMaxN = 2048;  // longest axis handled by one LDS-resident line
constexpr int kU = 8;           // global loads batched per thread before their LDS stores

// Lines per block: the largest power of two whose padded LDS footprint
// B*(n+1) complex values stays within 64 KB, and at most `avail`.
template <typename T>
static int fft_batch(int n, int64_t avail)
{
  const int64_t E = 65536 / (int64_t)(2 * sizeof(T)) - n / 2;  // twiddles share the LDS
  int B = 1;
  while ((int64_t)(2 * B) * (n + 1) <= E && 2 * B <= avail) B *= 2;
  return B;
}

template <typename T>
static size_t fft_lds(int n, int B) { return ((size_t)B * (n + 1) + (size_t)n / 2) * 2 * sizeof(T); }

// The z pass also stages 3 doubles per line and the z table (2n doubles).
template <typename T>
static size_t fft_lds_z(int n, int B) { return fft_lds<T>(n, B) + (3 * (size_t)B + 2 * (size_t)n) * sizeof(double); }

template <typename T>
static int fft_batch_z(int n, int64_t avail)
{
  int B = fft_batch<T>(n, avail);
  while (B > 1 && fft_lds_z<T>(n, B) > 65536) B >>= 1;
  return B;
}

// One batch of B lines in LDS: s[b*(n+1) + t] (padded rows).  Twiddles
// w[j] = exp(-2 pi i j / n), j < n/2, are staged in LDS (tw).  Butterfly
// indices use shifts and masks (n is a power of two).

// Forward, decimation in frequency (natural order in, bit-reversed out):
// radix-4 stages (two radix-2 stages fused), then a final radix-2 stage when
// log2 n is odd.
template <typename T>
__device__ __forceinline__ void dif_fwd(cpx<T> *s, int B, int logn, const cpx<T> *tw, int tsh = 0)
{
  const int n = 1 << logn;
  int lg = logn;
  for (; lg >= 2; lg -= 2) {
    const int len = 1 << lg, lq = lg - 2, q = 1 << lq, step = n >> lg;
    const int nb4 = B << (logn - 2);
    for (int m = threadIdx.x; m < nb4; m += blockDim.x) {
      const int b = m >> (logn - 2), r = m & ((n >> 2) - 1);
      const int grp = r >> lq, k = r & (q - 1);
      const int i0 = b * (n + 1) + grp * len + k;
      const cpx<T> x0 = s[i0], x1 = s[i0 + q], x2 = s[i0 + 2 * q], x3 = s[i0 + 3 * q];
      const cpx<T> w1 = tw[(k * step) << tsh], w2 = tw[(2 * k * step) << tsh];
      const cpx<T> a0 = {x0.x + x2.x, x0.y + x2.y};
      const cpx<T> a1 = {x1.x + x3.x, x1.y + x3.y};
      const cpx<T> d02 = {x0.x - x2.x, x0.y - x2.y};
      const cpx<T> d13 = {x1.x - x3.x, x1.y - x3.y};
      const cpx<T> a2 = cmul<T>(d02, w1);
      const cpx<T> a3 = cmul<T>({d13.y, -d13.x}, w1);  // (x1 - x3) W^k (-i)
      s[i0] = {a0.x + a1.x, a0.y + a1.y};
      s[i0 + q] = cmul<T>({a0.x - a1.x, a0.y - a1.y}, w2);
      s[i0 + 2 * q] = {a2.x + a3.x, a2.y + a3.y};
      s[i0 + 3 * q] = cmul<T>({a2.x - a3.x, a2.y - a3.y}, w2);
    }
    __syncthreads();
  }
  if (lg == 1) {
    const int nb2 = B << (logn - 1);
    for (int m = threadIdx.x; m < nb2; m += blockDim.x) {
      const int b = m >> (logn - 1), r = m & ((n >> 1) - 1);
      const int i0 = b * (n + 1) + 2 * r;
      const cpx<T> u = s[i0], v = s[i0 + 1];
      s[i0] = {u.x + v.x, u.y + v.y};
      s[i0 + 1] = {u.x - v.x, u.y - v.y};
    }
    __syncthreads();
  }
}

// Inverse, decimation in time (bit-reversed in, natural order out, unscaled):
// a radix-2 first stage when log2 n is odd, then radix-4 stages.
template <typename T>
__device__ __forceinline__ void dit_inv(cpx<T> *s, int B, int logn, const cpx<T> *tw, int tsh = 0)
{
  const int n = 1 << logn;
  int lg = 0;
  if (logn & 1) {
    const int nb2 = B << (logn - 1);
    for (int m = threadIdx.x; m < nb2; m += blockDim.x) {
      const int b = m >> (logn - 1), r = m & ((n >> 1) - 1);
      const int i0 = b * (n + 1) + 2 * r;
      const cpx<T> u = s[i0], v = s[i0 + 1];
      s[i0] = {u.x + v.x, u.y + v.y};
      s[i0 + 1] = {u.x - v.x, u.y - v.y};
    }
    __syncthreads();
    lg = 1;
  }
  for (; lg + 2 <= logn; lg += 2) {
    const int len = 1 << (lg + 2), lq = lg, q = 1 << lq, step = n >> (lg + 2);
    const int nb4 = B << (logn - 2);
    for (int m = threadIdx.x; m < nb4; m += blockDim.x) {
      const int b = m >> (logn - 2), r = m & ((n >> 2) - 1);
      const int grp = r >> lq, k = r & (q - 1);
      const int i0 = b * (n + 1) + grp * len + k;
      const cpx<T> x0 = s[i0], x1 = s[i0 + q], x2 = s[i0 + 2 * q], x3 = s[i0 + 3 * q];
      const cpx<T> t1 = tw[(k * step) << tsh], t2 = tw[(2 * k * step) << tsh];
      const cpx<T> W = {t1.x, -t1.y}, V = {t2.x, -t2.y};  // conjugates: +2 pi i
      const cpx<T> v1 = cmul<T>(x1, V), v3 = cmul<T>(x3, V);
      const cpx<T> a0 = {x0.x + v1.x, x0.y + v1.y}, a1 = {x0.x - v1.x, x0.y - v1.y};
      const cpx<T> a2 = {x2.x + v3.x, x2.y + v3.y}, a3 = {x2.x - v3.x, x2.y - v3.y};
      const cpx<T> b2 = cmul<T>(a2, W);
      const cpx<T> wa3 = cmul<T>(a3, W);
      const cpx<T> b3 = {-wa3.y, wa3.x};  // a3 W^k (+i)
      s[i0] = {a0.x + b2.x, a0.y + b2.y};
      s[i0 + 2 * q] = {a0.x - b2.x, a0.y - b2.y};
      s[i0 + q] = {a1.x + b3.x, a1.y + b3.y};
      s[i0 + 3 * q] = {a1.x - b3.x, a1.y - b3.y};
    }
    __syncthreads();
  }
}

template <typename T>
__device__ __forceinline__ void load_twiddles(cpx<T> *tws, const double *tw, int n)
{
  for (int j = threadIdx.x; j < (n >> 1); j += blockDim.x) tws[j] = {(T)tw[2 * j], (T)tw[2 * j + 1]};
}

// Pass along x (contiguous real lines of length n = 2h) as half-length
// complex transforms.  MODE 0 (forward): z_t = r_2t + i r_2t+1, Z = FFT_h(z),
// then X_k = E_k + W_n^k O_k with E_k = (Z_k + conj Z_h-k)/2,
// O_k = (Z_k - conj Z_h-k)/(2i), written for k < h in natural order; X_0 and
// X_h are real and share slot 0 as X_0 + i X_h.  MODE 1 (inverse): the exact
// reverse (E_k = X_k + conj X_h-k, O_k = (X_k - conj X_h-k) W_n^-k,
// Z = E + i O, z = IFFT_h(Z), r_2t + i r_2t+1 = z_t), which returns n times the
// input like an unscaled length-n inverse; the real result is multiplied by
// `scale`.  LDS: twiddles W_n^j (j < h; the half-length FFT uses every other
// one) and B padded lines of h + 1 complex values.
template <typename T>
static int rfft_batch(int h, int64_t avail)
{
  const int64_t E = 65536 / (int64_t)(2 * sizeof(T)) - h;
  int B = 1;
  while ((int64_t)(2 * B) * (h + 1) <= E && 2 * B <= avail) B *= 2;
  return B;
}

template <typename T>
static size_t rfft_lds(int h, int B) { return ((size_t)B * (h + 1) + (size_t)h) * 2 * sizeof(T); }

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_rfft_x(int logn, int nlines, int B, const T *__restrict__ rin,
                                                cpx<T> *__restrict__ buf, T *__restrict__ rout,
                                                const double *__restrict__ tw, double scale, int ly, int64_t plane)
{
  // half spectrum of line l = iy + ny*iz at iy*h + iz*plane
  auto at = [&](int64_t l) { return (l & ((1 << ly) - 1)) * (int64_t)(1 << (logn - 1)) + (l >> ly) * plane; };
  extern __shared__ unsigned char smem_raw[];
  const int n = 1 << logn, lh = logn - 1, h = n >> 1;
  cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
  cpx<T> *s = tws + h;
  load_twiddles<T>(tws, tw, n);
  const int64_t l0 = (int64_t)blockIdx.x * B;
  const int nb = (int)((nlines - l0) < B ? (nlines - l0) : B);
  const cpx<T> *rin2 = reinterpret_cast<const cpx<T> *>(rin);
  if (MODE == 0) {
    // kU independent loads in flight per thread before the LDS stores
    for (int e0 = threadIdx.x; e0 < (nb << lh); e0 += kU * 256) {
      cpx<T> v[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int e = e0 + 256 * u;
        if (e < (nb << lh)) v[u] = rin2[(l0 + (e >> lh)) * h + (e & (h - 1))];
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int e = e0 + 256 * u;
        if (e < (nb << lh)) s[(e >> lh) * (h + 1) + (e & (h - 1))] = v[u];
      }
    }
    __syncthreads();
    dif_fwd<T>(s, nb, lh, tws, 1);
    for (int e = threadIdx.x; e < (nb << lh); e += blockDim.x) {
      const int b = e >> lh, k = e & (h - 1);
      const cpx<T> Z = s[b * (h + 1) + (lh ? bitrev(k, lh) : 0)];
      cpx<T> X;
      if (k == 0) {
        X = {Z.x + Z.y, Z.x - Z.y};
      } else {
        const cpx<T> M = s[b * (h + 1) + bitrev(h - k, lh)];
        const cpx<T> E = {(T)0.5 * (Z.x + M.x), (T)0.5 * (Z.y - M.y)};
        const cpx<T> O = {(T)0.5 * (Z.y + M.y), (T)0.5 * (M.x - Z.x)};
        const cpx<T> WO = cmul<T>(O, tws[k]);
        X = {E.x + WO.x, E.y + WO.y};
      }
      buf[at(l0 + b) + k] = X;
    }
  } else {
    __syncthreads();  // the unpacking below reads twiddles loaded by other threads
    for (int e0 = threadIdx.x; e0 < (nb << lh); e0 += kU * 256) {
      cpx<T> X[kU], M[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int e = e0 + 256 * u;
        if (e < (nb << lh)) {
          const int b = e >> lh, k = e & (h - 1);
          const int64_t o = at(l0 + b);
          X[u] = buf[o + k];
          M[u] = buf[o + ((h - k) & (h - 1))];
        }
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const int e = e0 + 256 * u;
        if (e < (nb << lh)) {
          const int b = e >> lh, k = e & (h - 1);
          cpx<T> Z;
          if (k == 0) {
            Z = {X[u].x + X[u].y, X[u].x - X[u].y};
          } else {
            const cpx<T> E = {X[u].x + M[u].x, X[u].y - M[u].y};
            const cpx<T> D = {X[u].x - M[u].x, X[u].y + M[u].y};
            const cpx<T> w = tws[k];
            const cpx<T> O = cmul<T>(D, {w.x, -w.y});
            Z = {E.x - O.y, E.y + O.x};  // E + i O
          }
          s[b * (h + 1) + (lh ? bitrev(k, lh) : 0)] = Z;
        }
      }
    }
    __syncthreads();
    dit_inv<T>(s, nb, lh, tws, 1);
    cpx<T> *rout2 = reinterpret_cast<cpx<T> *>(rout);
    for (int e = threadIdx.x; e < (nb << lh); e += blockDim.x) {
      const int b = e >> lh, t = e & (h - 1);
      const cpx<T> z = s[b * (h + 1) + t];
      rout2[(l0 + b) * h + t] = {(T)((double)z.x * scale), (T)((double)z.y * scale)};
    }
  }
}

// Symbol of the constant-coefficient operator in closed form:
// lam = A0 + sum_a F_a c_a + sum_{a<b} (P_ab c_a c_b - M_ab s_a s_b),
// c_a = cos(theta_a), s_a = sin(theta_a) (faces and edges paired by a_o = a_-o).
struct Symbol {
  double A0, F[3], P[3], M[3];  // pairs ordered xy, xz, yz
  int nx, lx, ly, lz;
  int ynat;                     // y frequencies of the spectrum in natural (1) or bit-reversed (0) order
  double inv_scale;             // 1 / (nx ny nz)
};

__device__ __forceinline__ double symbol_at(const Symbol &sym, int kx, int ky, int kz, const double *__restrict__ trx,
                                            const double *__restrict__ try_, const double *__restrict__ trz)
{
  const double cx = trx[2 * kx], sx = trx[2 * kx + 1];
  const double cy = try_[2 * ky], sy = try_[2 * ky + 1];
  const double cz = trz[2 * kz], sz = trz[2 * kz + 1];
  return sym.A0 + sym.F[0] * cx + sym.F[1] * cy + sym.F[2] * cz + (sym.P[0] * cx * cy - sym.M[0] * sx * sy) +
         (sym.P[1] * cx * cz - sym.M[1] * sx * sz) + (sym.P[2] * cy * cz - sym.M[2] * sy * sz);
}

// inv_scale / lam, 0 for the null mode
__device__ __forceinline__ double symbol_factor(const Symbol &sym, int kx, int ky, int kz, const double *trx,
                                                const double *try_, const double *trz)
{
  if (kx == 0 && ky == 0 && kz == 0) return 0.0;
  const double lam = symbol_at(sym, kx, ky, kz, trx, try_, trz);
  return lam == 0.0 ? 0.0 : sym.inv_scale / lam;
}

// Pass along a strided axis: line (a, c) at base a + c*cstride, element t at
// + t*estride; a block holds B consecutive a.  MODE 0 forward, 1 inverse,
// 2 forward + symbol division + inverse (z axis) for the half-spectrum lines
// a = kx + (nx/2) * py with kx >= 1 in natural order (slot 0, which packs
// kx = 0 and nx/2, is left to k_fft_col0) and py the bit-reversed y
// frequency.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_fft_s(int logn, int64_t A, int64_t estride, int64_t cstride, int B,
                                               cpx<T> *__restrict__ buf, const double *__restrict__ tw, Symbol sym,
                                               const double *__restrict__ trx, const double *__restrict__ try_,
                                               const double *__restrict__ trz)
{
  extern __shared__ unsigned char smem_raw[];
  const int n = 1 << logn;
  cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
  cpx<T> *s = tws + (n >> 1);
  load_twiddles<T>(tws, tw, n);
  const int64_t blocks_per_c = A / B;
  const int64_t c = blockIdx.x / blocks_per_c;
  const int64_t a0 = (blockIdx.x % blocks_per_c) * B;
  const int64_t base = a0 + c * cstride;
  const int hx = sym.nx >> 1;
  const int lb = __ffs(B) - 1;  // B is a power of two
  // in MODE 2 the lines of the packed slot (kx = 0) are not touched
  auto live = [&](int b) { return MODE != 2 || ((a0 + b) & (hx - 1)) != 0; };
  for (int e0 = threadIdx.x; e0 < B * n; e0 += kU * 256) {
    cpx<T> v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int e = e0 + 256 * u;
      const int t = e >> lb, b = e & (B - 1);
      if (e < B * n && live(b)) v[u] = buf[base + b + (int64_t)t * estride];
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int e = e0 + 256 * u;
      const int t = e >> lb, b = e & (B - 1);
      if (e < B * n && live(b)) s[b * (n + 1) + t] = v[u];
    }
  }
  __syncthreads();
  if (MODE == 1) {
    dit_inv<T>(s, B, logn, tws);
  } else {
    dif_fwd<T>(s, B, logn, tws);
    if (MODE == 3) {
      // 2-D solve: line a is the bit-reversed x frequency, element t the
      // bit-reversed y frequency, kz = 0
      for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
        const int t = e >> lb, b = e & (B - 1);
        const int kx = bitrev((int)(a0 + b), sym.lx), ky = bitrev(t, sym.ly);
        const double f = symbol_factor(sym, kx, ky, 0, trx, try_, trz);
        cpx<T> &z = s[b * (n + 1) + t];
        z = {(T)((double)z.x * f), (T)((double)z.y * f)};
      }
      __syncthreads();
      dit_inv<T>(s, B, logn, tws);
    }
    if (MODE == 2) {
      // line a = kx + (nx/2)*py holds the natural x and bit-reversed y
      // frequencies, element t the bit-reversed z frequency.  Along a line the
      // symbol is la + lb cos(kz) - lc sin(kz) (the z terms of symbol_at
      // grouped): the three line constants and the z table are staged in LDS
      // after the lines (kx >= 1 here, so the null mode does not occur).
      double *lcn = reinterpret_cast<double *>(s + B * (n + 1));
      double *tz = lcn + 3 * B;
      for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const int64_t a = a0 + b;
        const int kx = (int)(a & (hx - 1));
        const int py = (int)(a >> (sym.lx - 1));
        const int ky = sym.ynat ? py : bitrev(py, sym.ly);
        const double cx = trx[2 * kx], sx = trx[2 * kx + 1], cy = try_[2 * ky], sy = try_[2 * ky + 1];
        lcn[3 * b] = sym.A0 + sym.F[0] * cx + sym.F[1] * cy + (sym.P[0] * cx * cy - sym.M[0] * sx * sy);
        lcn[3 * b + 1] = sym.F[2] + sym.P[1] * cx + sym.P[2] * cy;
        lcn[3 * b + 2] = sym.M[1] * sx + sym.M[2] * sy;
      }
      for (int t = threadIdx.x; t < 2 * n; t += blockDim.x) tz[t] = trz[t];
      __syncthreads();
      for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
        const int t = e >> lb, b = e & (B - 1);
        const int kz = bitrev(t, sym.lz);
        const double lam = lcn[3 * b] + lcn[3 * b + 1] * tz[2 * kz] - lcn[3 * b + 2] * tz[2 * kz + 1];
        const double f = lam == 0.0 ? 0.0 : sym.inv_scale / lam;
        cpx<T> &z = s[b * (n + 1) + t];
        z = {(T)((double)z.x * f), (T)((double)z.y * f)};
      }
      __syncthreads();
      dit_inv<T>(s, B, logn, tws);
    }
  }
  for (int e = threadIdx.x; e < B * n; e += blockDim.x) {
    const int t = e >> lb, b = e & (B - 1);
    if (live(b)) buf[base + b + (int64_t)t * estride] = s[b * (n + 1) + t];
  }
}

// The packed slot kx = 0 of the half spectrum holds X_0 + i X_h, both real in
// x, so after the y and z transforms Z = P + i Q with P = F(X_0), Q = F(X_h)
// Hermitian: P(k) = (Z(k) + conj Z(-k))/2, Q(k) = (Z(k) - conj Z(-k))/(2i).
// Block p takes the z-lines of bit-reversed y positions p and p' (ky' = -ky;
// p' == p for ky = 0, ny/2), transforms both, scales P by the symbol at
// kx = 0 and Q by the symbol at kx = nx/2, recombines and transforms back.
// Block p's share (all threads of the block): tws holds the n-point
// twiddles, s has room for 2 (n + 1) values.
template <typename T>
__device__ __forceinline__ void fft_col0_body(int p, int logn, int64_t estride, int hx, cpx<T> *__restrict__ buf,
                                              const cpx<T> *tws, cpx<T> *s, const Symbol &sym,
                                              const double *__restrict__ trx, const double *__restrict__ try_,
                                              const double *__restrict__ trz)
{
  const int n = 1 << logn;
  const int ny = 1 << sym.ly;
  const int ky = sym.ynat ? p : bitrev(p, sym.ly);
  const int pp = sym.ynat ? (ny - ky) & (ny - 1) : bitrev((ny - ky) & (ny - 1), sym.ly);
  if (pp < p) return;             // the pair is done by block pp
  const int nl = pp == p ? 1 : 2;
  const int64_t base[2] = {(int64_t)hx * p, (int64_t)hx * pp};
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    const int b = e >> logn, t = e & (n - 1);
    s[b * (n + 1) + t] = buf[base[b] + (int64_t)t * estride];
  }
  __syncthreads();
  dif_fwd<T>(s, nl, logn, tws);
  // every new value needs its partner's old one: compute into registers, then write
  constexpr int kPerThread = 2 * kFftMaxN / 256;
  cpx<T> nv[kPerThread];
#pragma unroll
  for (int q = 0; q < kPerThread; q++) {
    const int e = threadIdx.x + 256 * q;
    if (e >= nl * n) break;
    const int b = e >> logn, t = e & (n - 1);
    const int kyb = b == 0 ? ky : (ny - ky) & (ny - 1);
    const int kz = bitrev(t, sym.lz);
    const int tm = bitrev((n - kz) & (n - 1), sym.lz);
    const int bm = nl == 1 ? 0 : 1 - b;
    const cpx<T> Z = s[b * (n + 1) + t], M = s[bm * (n + 1) + tm];
    const double Px = 0.5 * ((double)Z.x + (double)M.x), Py = 0.5 * ((double)Z.y - (double)M.y);
    const double Qx = 0.5 * ((double)Z.y + (double)M.y), Qy = 0.5 * ((double)M.x - (double)Z.x);
    const double f0 = symbol_factor(sym, 0, kyb, kz, trx, try_, trz);
    const double fh = symbol_factor(sym, hx, kyb, kz, trx, try_, trz);
    // P f0 + i Q fh
    nv[q] = {(T)(Px * f0 - Qy * fh), (T)(Py * f0 + Qx * fh)};
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPerThread; q++) {
    const int e = threadIdx.x + 256 * q;
    if (e >= nl * n) break;
    s[(e >> logn) * (n + 1) + (e & (n - 1))] = nv[q];
  }
  __syncthreads();
  dit_inv<T>(s, nl, logn, tws);
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    const int b = e >> logn, t = e & (n - 1);
    buf[base[b] + (int64_t)t * estride] = s[b * (n + 1) + t];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fft_col0(int logn, int64_t estride, int hx, cpx<T> *__restrict__ buf,
                                                  const double *__restrict__ tw, Symbol sym,
                                                  const double *__restrict__ trx, const double *__restrict__ try_,
                                                  const double *__restrict__ trz)
{
  extern __shared__ unsigned char smem_raw[];
  const int n = 1 << logn;
  cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
  cpx<T> *s = tws + (n >> 1);     // 2 lines, padded
  load_twiddles<T>(tws, tw, n);
  fft_col0_body<T>((int)blockIdx.x, logn, estride, hx, buf, tws, s, sym, trx, try_, trz);
}

// ---- 256- and 512-point lines: register-resident four-step transforms -----------------
// A line of N = 16 R points (R = 16: 256, R = 32: 512) is held by 16 threads,
// thread j owning the elements j + 16 m (m < R) in registers.  Each thread
// transforms its R values (a length-R DFT in registers over m), multiplies by
// the twiddles W_N^(j k1), and one transpose through LDS hands thread j the
// 16 values of rows k1 = j + 16 c (c < R/16), whose 16-point DFTs over j
// give X[k1 + R k2].  That is element j + 16 (c + (R/16) k2): input and output
// use the same register pattern in natural order, so forward and inverse
// compose without permutations, and LDS carries each value twice per
// transform.  The LDS strides (values split into real and imaginary arrays)
// make the 64 lanes of a wave touch 64 distinct banks in both layouts.
constexpr int kQSk = 17;  // LDS stride between the k1 rows of a line
// line strides: strided passes (lanes: 16 lines x 4 threads), x pass (lanes:
// 16 threads x 4 line pairs); 4 and 16 modulo 64 respectively
template <int R> constexpr int q_sbs() { return R == 16 ? 324 : 580; }
template <int R> constexpr int q_sbx() { return R == 16 ? 272 : 592; }

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cpx<T> &a0, cpx<T> &a1, cpx<T> &a2, cpx<T> &a3)
{
  const cpx<T> t0 = {a0.x + a2.x, a0.y + a2.y}, t1 = {a0.x - a2.x, a0.y - a2.y};
  const cpx<T> t2 = {a1.x + a3.x, a1.y + a3.y}, t3 = {a1.x - a3.x, a1.y - a3.y};
  a0 = {t0.x + t2.x, t0.y + t2.y};
  a2 = {t0.x - t2.x, t0.y - t2.y};
  if (!INV) {  // X1 = t1 - i t3, X3 = t1 + i t3
    a1 = {t1.x + t3.y, t1.y - t3.x};
    a3 = {t1.x - t3.y, t1.y + t3.x};
  } else {
    a1 = {t1.x - t3.y, t1.y + t3.x};
    a3 = {t1.x + t3.y, t1.y - t3.x};
  }
}

// In-place 16-point DFT, natural order in and out: n = 4a + b, k = c + 4d.
template <typename T, bool INV>
__device__ __forceinline__ void dft16(cpx<T> (&v)[16])
{
  constexpr double c1 = 0.92387953251128674, s1 = 0.38268343236508978, h = 0.70710678118654752;  // cos, sin pi/8; sqrt 1/2
  // W_16^m = cos(2 pi m / 16) - i sin(2 pi m / 16) for the products m = b c
  constexpr double wr[10] = {1.0, c1, h, s1, 0.0, 0.0, -h, 0.0, 0.0, -c1};
  constexpr double wi[10] = {0.0, -s1, -h, -c1, -1.0, 0.0, -h, 0.0, 0.0, s1};
#pragma unroll
  for (int b = 0; b < 4; b++) dft4<T, INV>(v[b], v[4 + b], v[8 + b], v[12 + b]);
#pragma unroll
  for (int b = 1; b < 4; b++) {
#pragma unroll
    for (int c = 1; c < 4; c++) {
      const int m = b * c;
      const T x = (T)wr[m], y = INV ? (T)(-wi[m]) : (T)wi[m];
      v[4 * c + b] = cmul<T>(v[4 * c + b], {x, y});
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) dft4<T, INV>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
  // position 4c + d holds X[c + 4d]
  cpx<T> o[16];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int d = 0; d < 4; d++) o[c + 4 * d] = v[4 * c + d];
#pragma unroll
  for (int q = 0; q < 16; q++) v[q] = o[q];
}

// In-place R-point DFT (R = 16 or 32), natural order in and out; 32 points
// as two 16-point DFTs of the even and odd elements and one radix-2 stage.
template <typename T, bool INV, int R>
__device__ __forceinline__ void dftR(cpx<T> (&v)[R])
{
  if constexpr (R == 16) {
    dft16<T, INV>(v);
  } else {
    static_assert(R == 32, "16- or 32-point register DFTs");
    cpx<T> e[16], o[16];
#pragma unroll
    for (int q = 0; q < 16; q++) { e[q] = v[2 * q]; o[q] = v[2 * q + 1]; }
    dft16<T, INV>(e);
    dft16<T, INV>(o);
    // W_32^k = cos(2 pi k / 32) -+ i sin(2 pi k / 32)
    constexpr double wc[16] = {1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452, 0.7071067811865476, 0.5555702330196023, 0.38268343236508984, 0.19509032201612833, 0.0, -0.1950903220161282, -0.3826834323650897, -0.555570233019602, -0.7071067811865475, -0.8314696123025453, -0.9238795325112867, -0.9807852804032304};
    constexpr double ws[16] = {0.0, 0.19509032201612825, 0.3826834323650898, 0.5555702330196022, 0.7071067811865475, 0.8314696123025452, 0.9238795325112867, 0.9807852804032304, 1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025455, 0.7071067811865476, 0.5555702330196022, 0.3826834323650899, 0.1950903220161286};
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const T wr = (T)wc[k], wi = INV ? (T)ws[k] : (T)(-ws[k]);
      const cpx<T> t = cmul<T>(o[k], {wr, wi});
      v[k] = {e[k].x + t.x, e[k].y + t.y};
      v[k + 16] = {e[k].x - t.x, e[k].y - t.y};
    }
  }
}

// The four-step transform of the line held by this thread's group (thread j
// of 16; its LDS rows start at gbase).  tws: W_N^m, m < N/2 (forward).  Every
// thread of the block calls it (two barriers).
template <typename T, bool INV, int R>
__device__ __forceinline__ void fft_line(cpx<T> (&v)[R], int j, T *lre, T *lim, int gbase, const cpx<T> *tws)
{
  constexpr int N = 16 * R, H = N / 2;
  dftR<T, INV, R>(v);
#pragma unroll
  for (int k1 = 1; k1 < R; k1++) {
    const int m = j * k1;  // < N: W^m = -W^(m-N/2) past N/2
    cpx<T> w = tws[m & (H - 1)];
    if (m & H) w = {-w.x, -w.y};
    if (INV) w.y = -w.y;
    v[k1] = cmul<T>(v[k1], w);
  }
#pragma unroll
  for (int k1 = 0; k1 < R; k1++) {
    lre[gbase + k1 * kQSk + j] = v[k1].x;
    lim[gbase + k1 * kQSk + j] = v[k1].y;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < R / 16; c++) {
    cpx<T> u[16];
    const int row = gbase + (j + 16 * c) * kQSk;
#pragma unroll
    for (int n2 = 0; n2 < 16; n2++) u[n2] = {lre[row + n2], lim[row + n2]};
    dft16<T, INV>(u);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) v[c + (R / 16) * k2] = u[k2];
  }
  __syncthreads();
}

// Strided N-point pass: lines a = a0 .. a0+15 (16 per block) of the set c,
// line (a, c) at a + c*cstride, element t at + t*estride; thread (b, j) =
// (threadIdx & 15, threadIdx >> 4).  MODE 0 forward, 1 inverse, 2 forward +
// symbol division + inverse (z axis, lines a = kx + (nx/2)*py; the packed
// kx = 0 slot is left to k_fft_col0).
template <typename T, int MODE, int R>
__global__ __launch_bounds__(256) void k_fftq_s(int64_t A, int64_t estride, int64_t cstride,
                                                cpx<T> *__restrict__ buf, const double *__restrict__ tw, Symbol sym,
                                                const double *__restrict__ trx, const double *__restrict__ try_,
                                                const double *__restrict__ trz)
{
  constexpr int N = 16 * R, SB = q_sbs<R>();
  __shared__ T lre[16 * SB], lim[16 * SB];
  __shared__ cpx<T> tws[N / 2];
  load_twiddles<T>(tws, tw, N);
  __syncthreads();
  if constexpr (MODE == 2) {
    // blocks past the lines take the packed kx = 0 column (k_fft_col0's work,
    // the z lines the others leave), so it needs no launch of its own; lre is
    // their scratch
    static_assert(2 * (N + 1) * 2 <= 16 * SB, "column-0 scratch");
    const int64_t nmain = A / 16;
    if ((int64_t)blockIdx.x >= nmain) {
      fft_col0_body<T>((int)(blockIdx.x - nmain), R == 16 ? 8 : 9, estride, sym.nx >> 1, buf, tws,
                       reinterpret_cast<cpx<T> *>(lre), sym, trx, try_, trz);
      return;
    }
  }
  const int b = threadIdx.x & 15, j = threadIdx.x >> 4;
  const int64_t blocks_per_c = A / 16;
  const int64_t c = blockIdx.x / blocks_per_c, a = (blockIdx.x % blocks_per_c) * 16 + b;
  const int64_t base = a + c * cstride;
  const int hx = sym.nx >> 1;
  const bool live = MODE != 2 || (a & (hx - 1)) != 0;
  cpx<T> v[R];
#pragma unroll
  for (int m = 0; m < R; m++) v[m] = live ? buf[base + (int64_t)(j + 16 * m) * estride] : cpx<T>{(T)0, (T)0};
  fft_line<T, MODE == 1, R>(v, j, lre, lim, b * SB, tws);
  if (MODE == 2) {
    // symbol along the line: la + lb cos(kz) - lc sin(kz), kz = j + 16 m
    const int kx = (int)(a & (hx - 1));
    const int py = (int)(a >> (sym.lx - 1));
    const int ky = sym.ynat ? py : bitrev(py, sym.ly);
    const double cx = trx[2 * kx], sx = trx[2 * kx + 1], cy = try_[2 * ky], sy = try_[2 * ky + 1];
    const double la = sym.A0 + sym.F[0] * cx + sym.F[1] * cy + (sym.P[0] * cx * cy - sym.M[0] * sx * sy);
    const double lb = sym.F[2] + sym.P[1] * cx + sym.P[2] * cy;
    const double lc = sym.M[1] * sx + sym.M[2] * sy;
#pragma unroll
    for (int m = 0; m < R; m++) {
      const int kz = j + 16 * m;
      // lambda in FP64 (it cancels near k = 0); the FP32 pass then takes its
      // reciprocal in FP32, the FP64 pass divides exactly
      const double lam = la + lb * trz[2 * kz] - lc * trz[2 * kz + 1];
      if constexpr (sizeof(T) == 4) {
        const float f = lam == 0.0 ? 0.0f : __builtin_amdgcn_rcpf((float)lam) * (float)sym.inv_scale;
        v[m] = {v[m].x * f, v[m].y * f};
      } else {
        const double f = lam == 0.0 ? 0.0 : sym.inv_scale / lam;
        v[m] = {(T)((double)v[m].x * f), (T)((double)v[m].y * f)};
      }
    }
    fft_line<T, true, R>(v, j, lre, lim, b * SB, tws);
  }
  if (live) {
#pragma unroll
    for (int m = 0; m < R; m++) buf[base + (int64_t)(j + 16 * m) * estride] = v[m];
  }
}

// x pass for nx = N, two real lines per 16-thread group (32 lines per
// block): z = x_a + i x_b is transformed once, and X_a = (Z_k + conj Z_-k)/2,
// X_b = (Z_k - conj Z_-k)/(2i) are written for k < N/2 (slot 0 packs X_0 and
// X_(N/2) as X_0 + i X_(N/2)), to the same layout as k_rfft_x.  MODE 1 is the
// exact reverse: Z_k = X_a + i X_b over the Hermitian extension, inverse
// transform, x_a = Re, x_b = Im (N times the input), times scale.
template <typename T, int MODE, int R>
__global__ __launch_bounds__(256) void k_rfftq_x(int64_t nlines, const T *__restrict__ rin,
                                                 cpx<T> *__restrict__ buf, T *__restrict__ rout,
                                                 const double *__restrict__ tw, double scale, int ly, int64_t plane)
{
  constexpr int N = 16 * R, H = N / 2, SB = q_sbx<R>();
  __shared__ T lre[16 * SB], lim[16 * SB];
  __shared__ cpx<T> tws[H];
  load_twiddles<T>(tws, tw, N);
  __syncthreads();
  const int j = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int64_t la = (int64_t)blockIdx.x * 32 + 2 * g, lb = la + 1;
  const bool ok = lb < nlines;
  auto at = [&](int64_t l) { return (l & ((1 << ly) - 1)) * (int64_t)H + (l >> ly) * plane; };
  const int gb = g * SB;
  cpx<T> v[R];
  if (MODE == 0) {
#pragma unroll
    for (int m = 0; m < R; m++)
      v[m] = ok ? cpx<T>{rin[la * N + j + 16 * m], rin[lb * N + j + 16 * m]} : cpx<T>{(T)0, (T)0};
    fft_line<T, false, R>(v, j, lre, lim, gb, tws);
    // Z in natural order in LDS, for the partner values Z_-k
#pragma unroll
    for (int m = 0; m < R; m++) {
      lre[gb + j + 16 * m] = v[m].x;
      lim[gb + j + 16 * m] = v[m].y;
    }
    __syncthreads();
    if (ok) {
      const int64_t oa = at(la), ob = at(lb);
#pragma unroll
      for (int m = 0; m < R / 2; m++) {
        const int k = j + 16 * m;
        const T zx = lre[gb + k], zy = lim[gb + k];
        cpx<T> xa, xb;
        if (k == 0) {
          xa = {zx, lre[gb + H]};
          xb = {zy, lim[gb + H]};
        } else {
          const T mx = lre[gb + N - k], my = lim[gb + N - k];
          xa = {(T)0.5 * (zx + mx), (T)0.5 * (zy - my)};
          xb = {(T)0.5 * (zy + my), (T)0.5 * (mx - zx)};
        }
        buf[oa + k] = xa;
        buf[ob + k] = xb;
      }
    }
  } else {
    const int64_t oa = ok ? at(la) : 0, ob = ok ? at(lb) : 0;
    // each half-spectrum value is read once: A_t and B_t (t < N/2) staged in
    // LDS, then Z_t = A_t + i B_t and its Hermitian partner Z_(N-t) built from
    // there (Z_0 and Z_(N/2) from the packed slot 0)
    T *are = lre + gb, *aim = lim + gb, *bre = are + H, *bim = aim + H;
#pragma unroll
    for (int m = 0; m < R / 2; m++) {
      const int t = j + 16 * m;
      const cpx<T> A = ok ? buf[oa + t] : cpx<T>{(T)0, (T)0}, B = ok ? buf[ob + t] : cpx<T>{(T)0, (T)0};
      are[t] = A.x; aim[t] = A.y; bre[t] = B.x; bim[t] = B.y;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < R; m++) {
      const int t = j + 16 * m;
      cpx<T> z;
      if (t == 0) z = {are[0], bre[0]};
      else if (t == H) z = {aim[0], bim[0]};
      else if (t < H) z = {are[t] - bim[t], aim[t] + bre[t]};
      else z = {are[N - t] + bim[N - t], bre[N - t] - aim[N - t]};
      v[m] = z;
    }
    __syncthreads();
    fft_line<T, true, R>(v, j, lre, lim, gb, tws);
    if (ok) {
#pragma unroll
      for (int m = 0; m < R; m++) {
        rout[la * N + j + 16 * m] = (T)((double)v[m].x * scale);
        rout[lb * N + j + 16 * m] = (T)((double)v[m].y * scale);
      }
    }
  }
}

static bool pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

// Register radix R of the four-step kernels for an axis of n points: 16 for
// 256, 32 for 512 in FP32 (the FP64 512-point tiles would not fit LDS twice
// per CU); 0: the radix-4 LDS passes.
template <typename T> static int q_radix(int n) { return n == 256 ? 16 : (n == 512 && sizeof(T) == 4 ? 32 : 0); }

// Complex elements between z planes of the half spectrum: (nx/2) ny plus
// 256 bytes
template <typename T>
static int64_t spec_plane(int nx, int ny) { return (int64_t)(nx / 2) * ny + 256 / (int64_t)sizeof(cpx<T>); }
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

static Symbol make_symbol(const chy_ctx *ctx, const double v19[19])
{
  const int nx = ctx->g.nx, ny = ctx->g.ny, nz = ctx->g.nz;
  Symbol sym;
  sym.A0 = v19[18];
  sym.F[0] = 2.0 * v19[0]; sym.F[1] = 2.0 * v19[2]; sym.F[2] = 2.0 * v19[4];
  sym.P[0] = 2.0 * (v19[6] + v19[7]);   sym.M[0] = 2.0 * (v19[6] - v19[7]);
  sym.P[1] = 2.0 * (v19[10] + v19[11]); sym.M[1] = 2.0 * (v19[10] - v19[11]);
  sym.P[2] = 2.0 * (v19[14] + v19[15]); sym.M[2] = 2.0 * (v19[14] - v19[15]);
  sym.nx = nx;
  sym.lx = ilog2(nx); sym.ly = ilog2(ny); sym.lz = ilog2(nz);
  sym.inv_scale = 1.0 / ((double)nx * ny * nz);
  sym.ynat = 0;
  return sym;
}

template <typename T>
int fft_pc_apply(chy_ctx *ctx, const double v19[19], const T *r, T *z)
{
  const int nx = ctx->g.nx, ny = ctx->g.ny, nz = ctx->g.nz;
  cpx<T> *buf = (cpx<T> *)ctx->fft_buf;
  // symbol coefficients from the uniform-density row (column order of
  // lhs.c:11-104: faces 0..5, xy edges 6..9, xz 10..13, yz 14..17, centre 18)
  Symbol sym = make_symbol(ctx, v19);
  const int hx = nx / 2;
  // z planes of the half spectrum are spaced P = (nx/2) ny + padding apart, so
  // that the rows of a z line do not all fall at one power-of-two stride
  const int64_t P = spec_plane<T>(nx, ny);
  // 256-point axes (and 512-point ones in FP32) use the four-step kernels
  // (16 lines per block); their y pass leaves the y frequencies in natural
  // order.  xq, yq, zq: the register radix R of each axis (0: radix-4 passes)
  const int xq = q_radix<T>(nx), yq = hx % 16 == 0 ? q_radix<T>(ny) : 0,
            zq = ((int64_t)hx * ny) % 16 == 0 ? q_radix<T>(nz) : 0;
  sym.ynat = yq ? 1 : 0;
  const int nlines = ny * nz;
  const unsigned nxb = (unsigned)((nlines + 31) / 32), nyb = (unsigned)((hx / 16) * nz);
#define RFFTQ(MODE, RIN, ROUT)                                                                                        \
  do {                                                                                                             \
    if (xq == 16)                                                                                                  \
      hipLaunchKernelGGL((k_rfftq_x<T, MODE, 16>), nxb, 256, 0, ctx->stream, (int64_t)nlines, RIN, buf, ROUT,       \
                         ctx->tw[0], 1.0, ilog2(ny), P);                                                           \
    else if constexpr (sizeof(T) == 4)                                                                             \
      hipLaunchKernelGGL((k_rfftq_x<T, MODE, 32>), nxb, 256, 0, ctx->stream, (int64_t)nlines, RIN, buf, ROUT,       \
                         ctx->tw[0], 1.0, ilog2(ny), P);                                                           \
  } while (0)
#define FFTQ(MODE, RQ, NB, A_, ES, CS, TW)                                                                           \
  do {                                                                                                             \
    if (RQ == 16)                                                                                                  \
      hipLaunchKernelGGL((k_fftq_s<T, MODE, 16>), NB, 256, 0, ctx->stream, A_, ES, CS, buf, TW, sym, ctx->trig[0], \
                         ctx->trig[1], ctx->trig[2]);                                                              \
    else if constexpr (sizeof(T) == 4)                                                                             \
      hipLaunchKernelGGL((k_fftq_s<T, MODE, 32>), NB, 256, 0, ctx->stream, A_, ES, CS, buf, TW, sym, ctx->trig[0], \
                         ctx->trig[1], ctx->trig[2]);                                                              \
  } while (0)
  if (xq) {  // x forward: real lines -> half spectra (kx < nx/2, slot 0 packs kx = 0, nx/2)
    RFFTQ(0, r, (T *)nullptr);
  } else {
    const int B = rfft_batch<T>(hx, nlines);
    hipLaunchKernelGGL((k_rfft_x<T, 0>), (unsigned)((nlines + B - 1) / B), kBlock, rfft_lds<T>(hx, B), ctx->stream,
                       ilog2(nx), nlines, B, r, buf, (T *)nullptr, ctx->tw[0], 1.0, ilog2(ny), P);
  }
  // y forward: lines (kx, iz); A = nx/2 contiguous, c = iz with stride P
  if (yq) {
    FFTQ(0, yq, nyb, (int64_t)hx, (int64_t)hx, P, ctx->tw[1]);
  } else {
    const int n = ny, B = fft_batch<T>(n, hx);
    hipLaunchKernelGGL((k_fft_s<T, 0>), (unsigned)((hx / B) * nz), kBlock, fft_lds<T>(n, B), ctx->stream, ilog2(n),
                       (int64_t)hx, (int64_t)hx, P, B, buf, ctx->tw[1], sym, ctx->trig[0], ctx->trig[1],
                       ctx->trig[2]);
  }
  {  // z forward + symbol + inverse: lines a = kx + (nx/2)*py, estride P
    const int n = nz;
    const int64_t A = (int64_t)hx * ny;
    if (zq) {  // (ny more blocks take the kx = 0 column)
      FFTQ(2, zq, (unsigned)(A / 16 + ny), A, P, (int64_t)0, ctx->tw[2]);
    } else {
      const int B = fft_batch_z<T>(n, A);
      hipLaunchKernelGGL((k_fft_s<T, 2>), (unsigned)(A / B), kBlock, fft_lds_z<T>(n, B), ctx->stream, ilog2(n), A, P,
                         (int64_t)0, B, buf, ctx->tw[2], sym, ctx->trig[0], ctx->trig[1], ctx->trig[2]);
      const size_t lds = ((size_t)(n >> 1) + 2 * (size_t)(n + 1)) * 2 * sizeof(T);
      hipLaunchKernelGGL((k_fft_col0<T>), (unsigned)ny, kBlock, lds, ctx->stream, ilog2(n), P, hx, buf, ctx->tw[2],
                         sym, ctx->trig[0], ctx->trig[1], ctx->trig[2]);
    }
  }
  if (yq) {  // y inverse
    FFTQ(1, yq, nyb, (int64_t)hx, (int64_t)hx, P, ctx->tw[1]);
  } else {
    const int n = ny, B = fft_batch<T>(n, hx);
    hipLaunchKernelGGL((k_fft_s<T, 1>), (unsigned)((hx / B) * nz), kBlock, fft_lds<T>(n, B), ctx->stream, ilog2(n),
                       (int64_t)hx, (int64_t)hx, P, B, buf, ctx->tw[1], sym, ctx->trig[0], ctx->trig[1],
                       ctx->trig[2]);
  }
  if (xq) {  // x inverse: half spectra -> real lines
    RFFTQ(1, (const T *)nullptr, z);
  } else {
    const int B = rfft_batch<T>(hx, nlines);
    hipLaunchKernelGGL((k_rfft_x<T, 1>), (unsigned)((nlines + B - 1) / B), kBlock, rfft_lds<T>(hx, B), ctx->stream,
                       ilog2(nx), nlines, B, (const T *)nullptr, buf, z, ctx->tw[0], 1.0, ilog2(ny), P);
  }
#undef RFFTQ
#undef FFTQ
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, std::string("fft: ") + hipGetErrorString(e));
  return CHY_OK;
}

// ---- two-level preconditioner: kz = 0 correction -------------------------------------
// The 2-D inner GMRES runs on up to kNB2 blocks of 256 threads.  Every global
// sum is left as one partial per block; the kernel that needs it adds the
// partials itself (every block in the same fixed order, so all blocks agree
// bit for bit), which keeps one launch per Gram-Schmidt pass.
constexpr int kHm = kFft2Its;
constexpr int kNB2 = 256;
constexpr int kExtZ = 16;  // planes per block when the correction is added

static int nb2_of(int64_t nxy) { return (int)std::min<int64_t>(kNB2, (nxy + 255) / 256); }

// x-y fields in pc2_buf (T): C (9), V (kFft2Its + 1), Z (kFft2Its), w, and a
// complex work plane (2).
int64_t fft2_buf_elems(const chy_ctx *ctx)
{
  return (int64_t)(9 + (kFft2Its + 1) + kFft2Its + 1 + 2) * ctx->g.nx * ctx->g.ny;
}

// Double workspace in pc2_h: the z-chunk column sums part[zc][nxy], the
// stencil-block sums pl[nlb], then per-block partials of the inner GMRES:
// pn[j][nb] (|V_j|^2 before normalisation, j <= m), p1[j][q][nb] and
// p2[j][q][nb] (V_q . w in the two Gram-Schmidt passes of step j), and the
// least-squares coefficients y[m].
struct H2 {
  double *part, *pl, *pn, *p1, *p2, *y;
  int64_t nlb;
  int nb;
};

static H2 h2_layout(const chy_ctx *ctx, bool fp32)
{
  const int64_t nxy = (int64_t)ctx->g.nx * ctx->g.ny;
  H2 h;
  h.nb = nb2_of(nxy);
  h.nlb = lhs_zsum_blocks(ctx->g, fp32);
  h.part = ctx->pc2_h;
  h.pl = h.part + (int64_t)lhs_zchunks(ctx->g) * nxy;
  h.pn = h.pl + h.nlb;
  h.p1 = h.pn + (int64_t)(kHm + 1) * h.nb;
  h.p2 = h.p1 + (int64_t)kHm * kHm * h.nb;
  h.y = h.p2 + (int64_t)kHm * kHm * h.nb;
  return h;
}

int64_t fft2_h_elems(const chy_ctx *ctx)
{
  const int64_t nxy = (int64_t)ctx->g.nx * ctx->g.ny;
  const int nb = nb2_of(nxy);
  const int64_t nlb = std::max(lhs_zsum_blocks(ctx->g, false), lhs_zsum_blocks(ctx->g, true));
  return (int64_t)lhs_zchunks(ctx->g) * nxy + nlb + (int64_t)(kHm + 1) * nb +
         2 * (int64_t)kHm * kHm * nb + kHm;
}

// Block-wide sums of nv values per thread: out[q] (LDS) for q < nv.
template <int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], int nv, double *out)
{
  __shared__ double sh[16][NV];
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    if (q < nv) {
      double t = v[q];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
      if ((threadIdx.x & 63) == 0) sh[w][q] = t;
    }
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double t = 0.0;
    for (int k = 0; k < nw; k++) t += sh[k][threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// out[q] = sum over b < nb of p[q * nb + b], q < nv (LDS; same value in every block)
template <int NV>
__device__ __forceinline__ void sum_partials(const double *__restrict__ p, int nb, int nv, double *out)
{
  double a[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) a[q] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < NV; q++)
      if (q < nv) a[q] += p[(int64_t)q * nb + b];
  }
  block_sums<NV>(a, nv, out);
}

// V0 = r2 - mean(r2), with r2[col] = (1/nz) sum_k (r - A z) from the z-chunk
// column sums; pn[0][blk] = the block's part of |V0|^2
template <typename T>
__global__ __launch_bounds__(256) void k2_r2(int64_t nxy, int nz, int nzc, const double *__restrict__ part,
                                             const double *__restrict__ pl, int64_t nlb, T *__restrict__ V0,
                                             double *__restrict__ pn)
{
  __shared__ double m[1], q2[1];
  double a[1] = {0.0};
  for (int64_t b = threadIdx.x; b < nlb; b += blockDim.x) a[0] += pl[b];
  block_sums<1>(a, 1, m);
  const double inz = 1.0 / (double)nz, mean = m[0] * inz / (double)nxy;
  a[0] = 0.0;
  for (int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; col < nxy; col += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int zc = 0; zc < nzc; zc++) s += part[(int64_t)zc * nxy + col];
    const T x = (T)(s * inz - mean);
    V0[col] = x;
    a[0] += (double)x * (double)x;
  }
  block_sums<1>(a, 1, q2);
  if (threadIdx.x == 0) pn[blockIdx.x] = q2[0];
}

// 2-D complex pass along x.  MODE 0: V_j (before normalisation, |V_j|^2 in
// the partials pn) is scaled to unit norm (zero when the norm is zero),
// written back, and transformed forward.  MODE 1: inverse transform, real
// part to io.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_fft2_x(int logn, int nlines, int B, T *__restrict__ io,
                                                cpx<T> *__restrict__ buf, const double *__restrict__ tw,
                                                const double *__restrict__ pn, int nb)
{
  extern __shared__ unsigned char smem_raw[];
  const int n = 1 << logn;
  cpx<T> *tws = reinterpret_cast<cpx<T> *>(smem_raw);
  cpx<T> *s = tws + (n >> 1);
  __shared__ double hn[1];
  double inv = 1.0;
  if (MODE == 0) {
    sum_partials<1>(pn, nb, 1, hn);
    const double h = sqrt(hn[0]);
    inv = h > 0.0 ? 1.0 / h : 0.0;
  }
  load_twiddles<T>(tws, tw, n);
  const int64_t l0 = (int64_t)blockIdx.x * B;
  const int nbl = (int)((nlines - l0) < B ? (nlines - l0) : B);
  for (int e = threadIdx.x; e < (nbl << logn); e += blockDim.x) {
    const int b = e >> logn, t = e & (n - 1);
    const int64_t a = (l0 + b) * n + t;
    if (MODE == 0) {
      const T v = (T)((double)io[a] * inv);
      io[a] = v;
      s[b * (n + 1) + t] = {v, (T)0};
    } else {
      s[b * (n + 1) + t] = buf[a];
    }
  }
  __syncthreads();
  if (MODE == 0) dif_fwd<T>(s, nbl, logn, tws);
  else dit_inv<T>(s, nbl, logn, tws);
  for (int e = threadIdx.x; e < (nbl << logn); e += blockDim.x) {
    const int b = e >> logn, t = e & (n - 1);
    const int64_t a = (l0 + b) * n + t;
    if (MODE == 0) buf[a] = s[b * (n + 1) + t];
    else io[a] = s[b * (n + 1) + t].x;
  }
}

// w = A2 Z_j with the z-averaged 9-point coefficients; p1[q][blk] = the
// block's part of V_q . w, q <= j
template <typename T>
__global__ __launch_bounds__(256) void k2_a2dots(int nx, int ny, int j, const T *__restrict__ C,
                                                 const T *__restrict__ Zj, const T *__restrict__ V,
                                                 T *__restrict__ w, double *__restrict__ p1, int nb)
{
  __shared__ double h[kHm];
  const int64_t nxy = (int64_t)nx * ny;
  double a[kHm];
#pragma unroll
  for (int q = 0; q < kHm; q++) a[q] = 0.0;
  for (int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; col < nxy; col += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(col % nx), jy = (int)(col / nx);
    const int ip = i + 1 == nx ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
    const int jp = jy + 1 == ny ? 0 : jy + 1, jm = jy == 0 ? ny - 1 : jy - 1;
    const int xs[3] = {im, i, ip}, ys[3] = {jm, jy, jp};
    double acc = 0.0;
#pragma unroll
    for (int o = 0; o < 9; o++) acc += (double)C[o * nxy + col] * (double)Zj[xs[o % 3] + (int64_t)nx * ys[o / 3]];
    const T wt = (T)acc;
    w[col] = wt;
#pragma unroll
    for (int q = 0; q < kHm; q++)
      if (q <= j) a[q] += (double)V[(int64_t)q * nxy + col] * (double)wt;
  }
  block_sums<kHm>(a, j + 1, h);
  if (threadIdx.x <= j) p1[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// One classical Gram-Schmidt pass of step j: h = the sums of pin,
// x = w - sum_q h_q V_q.  PASS 1: w = x, pout[q][blk] = parts of V_q . x.
// PASS 2: V_{j+1} = x (normalised by the next pass along x), pout[blk] = part
// of |x|^2.
template <typename T, int PASS>
__global__ __launch_bounds__(256) void k2_gs_pass(int64_t nxy, int j, T *__restrict__ V, T *__restrict__ w,
                                                  const double *__restrict__ pin, double *__restrict__ pout, int nb)
{
  __shared__ double h[kHm], o[kHm];
  sum_partials<kHm>(pin, nb, j + 1, h);
  double a[kHm];
#pragma unroll
  for (int q = 0; q < kHm; q++) a[q] = 0.0;
  T *Vn = V + (int64_t)(j + 1) * nxy;
  for (int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; col < nxy; col += (int64_t)gridDim.x * blockDim.x) {
    double x = (double)w[col];
#pragma unroll
    for (int q = 0; q < kHm; q++)
      if (q <= j) x -= h[q] * (double)V[(int64_t)q * nxy + col];
    const T xt = (T)x;
    if (PASS == 1) {
      w[col] = xt;
#pragma unroll
      for (int q = 0; q < kHm; q++)
        if (q <= j) a[q] += (double)V[(int64_t)q * nxy + col] * (double)xt;
    } else {
      Vn[col] = xt;
      a[0] += (double)xt * (double)xt;
    }
  }
  const int nv = PASS == 1 ? j + 1 : 1;
  block_sums<kHm>(a, nv, o);
  if (threadIdx.x < nv) pout[(int64_t)threadIdx.x * nb + blockIdx.x] = o[threadIdx.x];
}

// Hessenberg columns from the partials, Givens QR and back substitution
// (one block): y[q], q < kHm (zero past a breakdown).
// (Every sum is formed as sum_partials forms it, so the values are those the
// other kernels used.)
constexpr int kNSums = (kHm + 1) + kHm * (kHm + 1);
__global__ __launch_bounds__(256) void k2_solve(const double *__restrict__ pn, const double *__restrict__ p1,
                                                const double *__restrict__ p2, int nb, double *__restrict__ y)
{
  __shared__ double red[kNSums];
  double a[kNSums];
#pragma unroll
  for (int t = 0; t < kNSums; t++) a[t] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int t = 0;
#pragma unroll
    for (int j = 0; j <= kHm; j++) a[t++] += pn[(int64_t)j * nb + b];
#pragma unroll
    for (int j = 0; j < kHm; j++)
#pragma unroll
      for (int q = 0; q <= j; q++) {
        a[t++] += p1[(int64_t)(j * kHm + q) * nb + b];
        a[t++] += p2[(int64_t)(j * kHm + q) * nb + b];
      }
  }
  block_sums<kNSums>(a, kNSums, red);
  if (threadIdx.x != 0) return;
  double n2[kHm + 1], h1[kHm][kHm], h2[kHm][kHm];
  {
    int t = 0;
    for (int j = 0; j <= kHm; j++) n2[j] = red[t++];
    for (int j = 0; j < kHm; j++)
      for (int q = 0; q <= j; q++) {
        h1[j][q] = red[t++];
        h2[j][q] = red[t++];
      }
  }
  double H[kHm][kHm], g[kHm + 1], cs[kHm], sn[kHm], yy[kHm];
  for (int q = 0; q <= kHm; q++) g[q] = 0.0;
  for (int q = 0; q < kHm; q++) yy[q] = 0.0;
  const double beta = sqrt(n2[0]);
  int jm = 0;
  if (beta > 0.0) {
    g[0] = beta;
    for (int j = 0; j < kHm; j++) {
      double h[kHm + 1];
      for (int q = 0; q <= j; q++) h[q] = h1[j][q] + h2[j][q];
      h[j + 1] = sqrt(n2[j + 1]);
      const bool more = h[j + 1] > 0.0;
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
      for (int q = 0; q <= j; q++) H[q][j] = h[q];
      jm = j + 1;
      if (!more) break;
    }
    for (int q = jm - 1; q >= 0; q--) {
      double s = g[q];
      for (int r = q + 1; r < jm; r++) s -= H[q][r] * yy[r];
      yy[q] = H[q][q] != 0.0 ? s / H[q][q] : 0.0;
    }
  }
  for (int q = 0; q < kHm; q++) y[q] = yy[q];
}

// z[k * nxy + col] += sum_q y_q Z_q[col] over kExtZ planes per block
template <typename T>
__global__ __launch_bounds__(256) void k2_ext(int64_t nxy, int nz, const T *__restrict__ Z, const double *__restrict__ y,
                                              T *__restrict__ z)
{
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= nxy) return;
  double e = 0.0;
#pragma unroll
  for (int q = 0; q < kHm; q++) e += y[q] * (double)Z[(int64_t)q * nxy + col];
  const int k0 = blockIdx.y * kExtZ;
  T v[kExtZ];
#pragma unroll
  for (int u = 0; u < kExtZ; u++)
    if (k0 + u < nz) v[u] = z[(int64_t)(k0 + u) * nxy + col];
#pragma unroll
  for (int u = 0; u < kExtZ; u++)
    if (k0 + u < nz) z[(int64_t)(k0 + u) * nxy + col] = (T)((double)v[u] + e);
}

// Lines per block of a 2-D pass: at least 64 blocks when there are lines enough
template <typename T>
static int fft2_batch(int n, int64_t lines)
{
  return fft_batch<T>(n, std::max<int64_t>(1, lines / 64));
}

template <typename T>
int fft2_pc_apply(chy_ctx *ctx, const double v19[19], const T *r, T *z, T *w3, bool zline, bool *zline_done)
{
  if (zline_done) *zline_done = false;
  TRY(fft_pc_apply<T>(ctx, v19, r, z));
  if (ctx->solve_transpose) return CHY_OK;  // the correction is built for A(n), not its transpose
  // the z-line stage fused with the correction: the column-sum pass also keeps
  // r - A(n) z in w3 for it, and e2 is added by the stage itself
  const bool fused = zline && w3 && ctx->zl_buf && zline_e2_applies(ctx);
  const int nx = ctx->g.nx, ny = ctx->g.ny, nz = ctx->g.nz;
  const int64_t nxy = (int64_t)nx * ny;
  T *C = (T *)ctx->pc2_buf, *V = C + 9 * nxy, *Z = V + (kFft2Its + 1) * nxy, *w = Z + kFft2Its * nxy;
  cpx<T> *cb = (cpx<T> *)(w + nxy);
  const H2 h = h2_layout(ctx, sizeof(T) == 4);
  const int nb = h.nb;
  // z-averaged residual of the first level, mean removed, into V0
  launch_lhs_zsum<T>(ctx, ctx->prm.lhs_type, (const T *)ctx->mom[0], z, r, h.part, h.pl, fused ? w3 : nullptr);
  hipLaunchKernelGGL((k2_r2<T>), nb, 256, 0, ctx->stream, nxy, nz, lhs_zchunks(ctx->g), (const double *)h.part,
                     (const double *)h.pl, h.nlb, V, h.pn);
  Symbol sym = make_symbol(ctx, v19);
  sym.inv_scale = 1.0 / (double)nxy;
  const int Bx = fft2_batch<T>(nx, ny), By = fft2_batch<T>(ny, nx);
  for (int j = 0; j < kFft2Its; j++) {
    T *Vj = V + (int64_t)j * nxy, *Zj = Z + (int64_t)j * nxy;
    // Zj = F2 Vj: 2-D FFT solve with the kz = 0 symbol
    hipLaunchKernelGGL((k_fft2_x<T, 0>), (unsigned)((ny + Bx - 1) / Bx), kBlock, fft_lds<T>(nx, Bx), ctx->stream,
                       ilog2(nx), ny, Bx, Vj, cb, ctx->tw[0], (const double *)(h.pn + (int64_t)j * nb), nb);
    hipLaunchKernelGGL((k_fft_s<T, 3>), (unsigned)(nx / By), kBlock, fft_lds<T>(ny, By), ctx->stream, ilog2(ny),
                       (int64_t)nx, (int64_t)nx, (int64_t)0, By, cb, ctx->tw[1], sym, ctx->trig[0], ctx->trig[1],
                       ctx->trig[2]);
    hipLaunchKernelGGL((k_fft2_x<T, 1>), (unsigned)((ny + Bx - 1) / Bx), kBlock, fft_lds<T>(nx, Bx), ctx->stream,
                       ilog2(nx), ny, Bx, Zj, cb, ctx->tw[0], (const double *)nullptr, 0);
    double *p1 = h.p1 + (int64_t)j * kHm * nb, *p2 = h.p2 + (int64_t)j * kHm * nb;
    hipLaunchKernelGGL((k2_a2dots<T>), nb, 256, 0, ctx->stream, nx, ny, j, (const T *)C, (const T *)Zj,
                       (const T *)V, w, p1, nb);
    hipLaunchKernelGGL((k2_gs_pass<T, 1>), nb, 256, 0, ctx->stream, nxy, j, V, w, (const double *)p1, p2, nb);
    hipLaunchKernelGGL((k2_gs_pass<T, 2>), nb, 256, 0, ctx->stream, nxy, j, V, w, (const double *)p2,
                       h.pn + (int64_t)(j + 1) * nb, nb);
  }
  hipLaunchKernelGGL(k2_solve, 1, 256, 0, ctx->stream, (const double *)h.pn, (const double *)h.p1,
                     (const double *)h.p2, nb, h.y);
  if (fused) {
    launch_zline_apply_e2<T>(ctx, (const T *)ctx->zl_buf, ctx->zl_col, w3, z, (const T *)ctx->mom[0], (const T *)Z,
                             (const double *)h.y);
    if (zline_done) *zline_done = true;
  } else {
    hipLaunchKernelGGL((k2_ext<T>), dim3((unsigned)((nxy + 255) / 256), (unsigned)((nz + kExtZ - 1) / kExtZ)), 256,
                       0, ctx->stream, nxy, nz, (const T *)Z, (const double *)h.y, z);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, std::string("fft2: ") + hipGetErrorString(e));
  return CHY_OK;
}

template int fft2_pc_apply<double>(chy_ctx *, const double *, const double *, double *, double *, bool, bool *);
template int fft2_pc_apply<float>(chy_ctx *, const double *, const float *, float *, float *, bool, bool *);

template int fft_pc_apply<double>(chy_ctx *, const double *, const double *, double *);
template int fft_pc_apply<float>(chy_ctx *, const double *, const float *, float *);

}  // namespace chy
