// chy_device.h -- device-side helpers shared by the gfx950 kernels:
// counter-based RNG, periodic cell arithmetic, CIC corner weights and the
// closed-form mean relative speed used to normalise collision probabilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chy {

// Grid descriptor passed by value to kernels.  Nodes are periodic and
// node-centred: node (i,j,k) at p0 + (i,j,k)*d, linear c = i + nx*(j + ny*k).
struct Grid {
  int nx, ny, nz;
  int64_t nc;
  double p0[3], p1[3], d[3], L[3];
  // ceil(2^64 / nx) and ceil(2^64 / ny): q = mulhi64(m, c) is c / n exactly for
  // every 32-bit c and n >= 2 (Lemire, Kaser and Kurz, "Faster remainder by
  // direct computation", 2019)
  uint64_t mx, my;
};

inline uint64_t fastdiv_magic(uint32_t n) { return ~0ull / n + 1; }

// cell -> (i, j, k) without runtime integer division
__device__ __forceinline__ void split_cell(const Grid &g, int c, int &i, int &j, int &k)
{
  const uint32_t q = (uint32_t)__umul64hi(g.mx, (uint64_t)(uint32_t)c);
  i = c - (int)q * g.nx;
  const uint32_t r = (uint32_t)__umul64hi(g.my, (uint64_t)q);
  j = (int)q - (int)r * g.ny;
  k = (int)r;
}

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Logical tile of this workgroup in a 3-D launch.  Workgroups are dealt to the
// 8 XCDs round-robin in linear order (x fastest); relabelling them so that XCD
// q takes the q-th contiguous eighth of the tiles keeps a tile's y and z
// neighbours, whose halo rows it reads, in the same L2.  Returns the logical
// linear index; (tx, ty, tz) are its coordinates.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t &tx, uint32_t &ty, uint32_t &tz)
{
  const uint32_t gx = gridDim.x, gy = gridDim.y, nb = gx * gy * gridDim.z;
  uint32_t b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  if ((nb & 7u) == 0) b = (b & 7u) * (nb >> 3) + (b >> 3);
  const uint32_t p = b / gx;
  tx = b - p * gx;
  ty = p % gy;
  tz = p / gy;
  return b;
}

__device__ __forceinline__ int64_t lin(const Grid &g, int i, int j, int k)
{
  return (int64_t)wrapi(i, g.nx) + (int64_t)g.nx * ((int64_t)wrapi(j, g.ny) + (int64_t)g.ny * wrapi(k, g.nz));
}

// ---- Philox4x32-10 (Salmon et al., SC'11) ------------------------------------
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
  }
  return c;
}

// 24-bit uniform in (0,1), exactly representable in FP32 and FP64.
template <typename T>
__device__ __forceinline__ T u24(uint32_t w) { return ((T)(w >> 8) + (T)0.5) * (T)(1.0 / 16777216.0); }

// RNG streams: counter = {id, step, stream, block}, key = seed.
enum : uint32_t { STREAM_INIT = 0, STREAM_COLLIDE_A = 1, STREAM_COLLIDE_B = 2, STREAM_WEIGHT = 3 };

// ---- periodic cells ----------------------------------------------------------
// Cell of an absolute coordinate, clamped to [0, n-1]; f = fraction in cell.
__device__ __forceinline__ void cell_coord(double x, double p0, double d, int n, int &i, double &f)
{
  const double xi = (x - p0) / d;
  int ii = (int)xi;
  if (ii > n - 1) ii = n - 1;
  if (ii < 0) ii = 0;
  i = ii;
  f = xi - (double)ii;
}

// Periodic wrap into [p0, p1); a coordinate more than one period away is
// first reduced with floor, non-finite values are returned unchanged.
__device__ __forceinline__ double wrap_pos(double x, double p0, double p1, double L)
{
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

// a*b + c: one fused multiply-add in FP32 (the mixed path, compared with the
// oracle to a tolerance); in FP64 a separate multiply and add, rounded as the
// oracle rounds them (the build turns off FMA contraction).
__device__ __forceinline__ float fmad(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return a * b + c; }

// Corner weights W_q = (wz*wy)*wx for q = dx + 2 dy + 4 dz.
template <typename T>
__device__ __forceinline__ void corner_weights(T fx, T fy, T fz, T W[8])
{
  const T wx[2] = {(T)1 - fx, fx}, wy[2] = {(T)1 - fy, fy}, wz[2] = {(T)1 - fz, fz};
#pragma unroll
  for (int q = 0; q < 8; q++) W[q] = (wz[q >> 2] * wy[(q >> 1) & 1]) * wx[q & 1];
}

// E|v - v_n| for v_n ~ N(vn0, vnT^2 I): mean of a non-central chi-3 variable,
// sigma [sqrt(2/pi) e^{-l^2/2} + (l + 1/l) erf(l/sqrt 2)], l = |v - vn0|/sigma.
__device__ __forceinline__ double mean_rel_speed(double vx, double vy, double vz,
                                                 double n0x, double n0y, double n0z, double vnT)
{
  const double ux = vx - n0x, uy = vy - n0y, uz = vz - n0z;
  const double mu = sqrt(ux * ux + uy * uy + uz * uz);
  if (!(vnT > 0.0)) return mu;
  const double lam = mu / vnT;
  const double c = 0.79788456080286535588;
  if (lam < 1e-3) return vnT * c * (2.0 + lam * lam / 3.0);
  return vnT * (c * exp(-0.5 * lam * lam) + (lam + 1.0 / lam) * erf(lam * 0.70710678118654752440));
}

__device__ __forceinline__ float mean_rel_speed_f(float vx, float vy, float vz, float n0x, float n0y, float n0z,
                                                  float vnT)
{
  const float ux = vx - n0x, uy = vy - n0y, uz = vz - n0z;
  const float mu = sqrtf(ux * ux + uy * uy + uz * uz);
  if (!(vnT > 0.f)) return mu;
  const float lam = mu / vnT;
  const float c = 0.79788456080286535588f;
  if (lam < 1e-2f) return vnT * c * (2.0f + lam * lam / 3.0f);
  return vnT * (c * expf(-0.5f * lam * lam) + (lam + 1.0f / lam) * erff(lam * 0.70710678118654752440f));
}

// ---- wave-level helpers (64-wide wavefronts) -----------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Run-length aggregation of integer atomics over lanes holding equal keys in
// consecutive lanes: returns, for each active lane, (head lane, run length).
// Inactive lanes must pass key = -1.
__device__ __forceinline__ void key_runs(int key, int &head, int &len)
{
  const int l = lane_id();
  const int prev = __shfl_up(key, 1, 64);
  const bool is_head = (l == 0) || (prev != key);
  const uint64_t heads = __ballot(is_head);
  // head of my run: highest set bit at or below l
  const uint64_t below = heads & ((l == 63) ? ~0ull : ((2ull << l) - 1));
  head = 63 - __clzll(below);
  const uint64_t above = heads & ~((l == 63) ? ~0ull : ((2ull << l) - 1));
  const int next = above ? __ffsll((unsigned long long)above) - 1 : 64;
  len = next - head;
}

}  // namespace chy
