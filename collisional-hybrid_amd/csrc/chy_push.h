// chy_push.h -- per-ion pieces of the push shared by the forward kernels and
// the adjoint: Boris constants, trilinear gather, Boris rotation, the
// collision decision and scattering kinematics, and the FP64 move.
#pragma once

#include "chy_internal.h"

namespace chy {

struct PushConst {
  double tau, t[3], s[3];          // Boris constants (particles.c:374-393)
  double dt, cdt[3];               // dt and dt/d per axis (mixed move)
  double mi, mn, viT, vnT, vn0[3];
  double Nc;
  uint32_t k0, k1, step;
};

inline PushConst push_const(const chy_ctx *ctx)
{
  const chy_params &p = ctx->prm;
  PushConst c;
  c.tau = 0.5 * (p.qi / p.mi) * p.dt;
  for (int a = 0; a < 3; a++) c.t[a] = c.tau * p.B0[a];
  const double tt = c.t[0] * c.t[0] + c.t[1] * c.t[1] + c.t[2] * c.t[2];
  for (int a = 0; a < 3; a++) c.s[a] = 2.0 * c.t[a] / (1 + tt);
  c.dt = p.dt;
  for (int a = 0; a < 3; a++) c.cdt[a] = p.dt / p.d[a];
  c.mi = p.mi; c.mn = p.mn; c.viT = p.viT_r; c.vnT = p.vnT;
  for (int a = 0; a < 3; a++) c.vn0[a] = p.vn0[a];
  c.Nc = 0.0;
  c.k0 = (uint32_t)p.seed;
  c.k1 = (uint32_t)(p.seed >> 32);
  c.step = (uint32_t)ctx->step;
  return c;
}

// The same constants in FP32 for the mixed kernels: passed as kernel
// arguments they sit in scalar registers, where the FP64 set would be
// converted per wave into vector registers.  Every conversion is the
// round-to-nearest one the FP32 arithmetic applied to the FP64 values.
struct PushConstF {
  float tau, t[3], s[3];
  float dt, cdt[3];
  float mi, mn, viT, vnT, vn0[3];
  float Nc;
  uint32_t k0, k1, step;
};

inline PushConstF push_const_f(const PushConst &c)
{
  PushConstF f;
  f.tau = (float)c.tau;
  for (int a = 0; a < 3; a++) {
    f.t[a] = (float)c.t[a]; f.s[a] = (float)c.s[a]; f.cdt[a] = (float)c.cdt[a]; f.vn0[a] = (float)c.vn0[a];
  }
  f.dt = (float)c.dt;
  f.mi = (float)c.mi; f.mn = (float)c.mn; f.viT = (float)c.viT; f.vnT = (float)c.vnT;
  f.Nc = (float)c.Nc;
  f.k0 = c.k0; f.k1 = c.k1; f.step = c.step;
  return f;
}

inline int64_t collision_count(const chy_params &p, int64_t Np)
{
  // particles.c:510-519
  const double fc = p.nui * p.dt;
  const double dvx = p.vi0[0] - p.vn0[0], dvy = p.vi0[1] - p.vn0[1], dvz = p.vi0[2] - p.vn0[2];
  const double vrm = 4.0 * p.viT_r + sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
  const double den = (p.vnT + p.vn0_r) * sqrt(p.mn / p.mi);
  if (!(den > 0.0)) return Np;
  const double x = ((double)Np * fc * vrm) / den;
  if (!(x < (double)Np)) return Np;
  if (!(x > 0.0)) return 0;
  return (int64_t)x;
}

// Arithmetic of the collision kinematics.  FP64 (the reference's
// arithmetic): IEEE square root, division, sine and cosine.  FP32 (the mixed
// path): the hardware square root, reciprocal and sine/cosine, each within a
// few ulp -- the neutral draw and the scattering are Monte Carlo samples whose
// FP32 rounding is far below their sampling noise -- with the square root
// clamped at zero against rounding that pushes 1 - cos^2 below it.
template <typename T> __device__ __forceinline__ T kin_sqrt(T x) { return sqrt(x); }
template <> __device__ __forceinline__ float kin_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(fmaxf(x, 0.0f)); }
template <typename T> __device__ __forceinline__ T kin_div(T a, T b) { return a / b; }
template <> __device__ __forceinline__ float kin_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename T> __device__ __forceinline__ T kin_cos(T a) { return cos(a); }
template <> __device__ __forceinline__ float kin_cos<float>(float a) { return __cosf(a); }
template <typename T> __device__ __forceinline__ T kin_sin(T a) { return sin(a); }
template <> __device__ __forceinline__ float kin_sin<float>(float a) { return __sinf(a); }

// ---- ion-neutral collisions (reference particles.c:529-611) ---------------------
// Neutral draw and Bernoulli decision from one Philox block: Box-Muller on the
// top 24 bits of the four words, the decision uniform from their low bytes.
// The decision is "collide iff u < Nc |v_r| / S" (probability clipped at 1).
template <typename T, typename PC>
__device__ __forceinline__ void neutral_draw(const PC &pc, uint32_t pid, T &vnx, T &vny, T &vnz, T &udec)
{
  const U4 a = philox4x32_10(U4{pid, pc.step, STREAM_COLLIDE_A, 0u}, pc.k0, pc.k1);
  const T u0 = u24<T>(a.x), u1 = u24<T>(a.y), u2 = u24<T>(a.z), u3 = u24<T>(a.w);
  udec = u24<T>(((a.x & 0xffu) | ((a.y & 0xffu) << 8) | ((a.z & 0xffu) << 16)) << 8);
  T r1, r2, c1, s1, c2;
  if (sizeof(T) == 4) {
    // hardware transcendentals: ample accuracy for a Monte Carlo neutral draw
    r1 = kin_sqrt<float>(-2.0f * __logf((float)u0));
    r2 = kin_sqrt<float>(-2.0f * __logf((float)u2));
    const float a1 = 6.28318530717958647692f * (float)u1, a2 = 6.28318530717958647692f * (float)u3;
    c1 = __cosf(a1); s1 = __sinf(a1); c2 = __cosf(a2);
  } else {
    r1 = sqrt(-2.0 * log((double)u0));
    r2 = sqrt(-2.0 * log((double)u2));
    const double a1 = 2.0 * M_PI * (double)u1, a2 = 2.0 * M_PI * (double)u3;
    c1 = cos(a1); s1 = sin(a1); c2 = cos(a2);
  }
  const T vnT = (T)pc.vnT;
  vnx = (T)pc.vn0[0] + vnT * (r1 * c1);
  vny = (T)pc.vn0[1] + vnT * (r1 * s1);
  vnz = (T)pc.vn0[2] + vnT * (r2 * c2);
}

// Collision probability Nc |v_r| / S (clipped at 1) for a neutral draw vn.
template <typename T, typename PC>
__device__ __forceinline__ T collide_prob_n(const PC &pc, T S, T vnx, T vny, T vnz, T vx, T vy, T vz, int &capped)
{
  const T vrx = vx - vnx, vry = vy - vny, vrz = vz - vnz;
  const T vrr = kin_sqrt<T>(vrx * vrx + vry * vry + vrz * vrz);
  T p = (S > (T)0) ? kin_div<T>((T)pc.Nc * vrr, S) : (T)0;
  capped = 0;
  if (p > (T)1) { p = (T)1; capped = 1; }
  return p;
}

// Collision decision from a neutral draw (vn, ud) already made.
template <typename T, typename PC>
__device__ __forceinline__ bool collide_decide_n(const PC &pc, T S, T vnx, T vny, T vnz, T ud, T vx, T vy,
                                                 T vz, int &capped)
{
  return ud < collide_prob_n<T>(pc, S, vnx, vny, vnz, vx, vy, vz, capped);
}

template <typename T, typename PC>
__device__ __forceinline__ bool collide_decide(const PC &pc, T S, uint32_t pid, T vx, T vy, T vz, int &capped)
{
  T vnx, vny, vnz, ud;
  neutral_draw<T>(pc, pid, vnx, vny, vnz, ud);
  const T vrx = vx - vnx, vry = vy - vny, vrz = vz - vnz;
  const T vrr = kin_sqrt<T>(vrx * vrx + vry * vry + vrz * vrz);
  T p = (S > (T)0) ? kin_div<T>((T)pc.Nc * vrr, S) : (T)0;
  capped = 0;
  if (p > (T)1) { p = (T)1; capped = 1; }
  return ud < p;
}

// Elastic scattering in the centre-of-mass frame, rotated into the frame of
// the relative velocity (particles.c:552-611).  Returns 1 accepted, 2 refused.
template <typename T, typename PC>
__device__ __forceinline__ int scatter_n(const PC &pc, uint32_t pid, T vnx, T vny, T vnz, T &vx, T &vy, T &vz);

template <typename T, typename PC>
__device__ __forceinline__ int scatter(const PC &pc, uint32_t pid, T &vx, T &vy, T &vz)
{
  T vnx, vny, vnz, ud;
  neutral_draw<T>(pc, pid, vnx, vny, vnz, ud);
  return scatter_n<T>(pc, pid, vnx, vny, vnz, vx, vy, vz);
}

// scatter with the neutral draw (vn) already made
template <typename T, typename PC>
__device__ __forceinline__ int scatter_n(const PC &pc, uint32_t pid, T vnx, T vny, T vnz, T &vx, T &vy, T &vz)
{
  const T vrx = vx - vnx, vry = vy - vny, vrz = vz - vnz;
  const T vrr = kin_sqrt<T>(vrx * vrx + vry * vry + vrz * vrz);
  const U4 b = philox4x32_10(U4{pid, pc.step, STREAM_COLLIDE_B, 0u}, pc.k0, pc.k1);
  const T twopi = (T)(2.0 * M_PI);
  const T mi = (T)pc.mi, mn = (T)pc.mn, M = mn + mi;
  const T vcx = kin_div<T>(vx * mi + vnx * mn, M);
  const T vcy = kin_div<T>(vy * mi + vny * mn, M);
  const T vcz = kin_div<T>(vz * mi + vnz * mn, M);
  const T ex = vx - vcx, ey = vy - vcy, ez = vz - vcz;
  const T vcr = kin_sqrt<T>(ex * ex + ey * ey + ez * ez);
  const T uz = (T)2.0 * u24<T>(b.x) - (T)1.0;
  const T uperp = kin_sqrt<T>((T)1.0 - uz * uz);
  const T uphi = twopi * u24<T>(b.y);
  const T ux = uperp * kin_cos<T>(uphi), uy = uperp * kin_sin<T>(uphi);
  const T costht = kin_div<T>(vrz, vrr);
  const T sintht = kin_sqrt<T>((T)1.0 - costht * costht);
  T cosphi = (T)1, sinphi = (T)0;
  if (sintht > (T)1e-12) {
    cosphi = kin_div<T>(vrx, vrr * sintht);
    sinphi = kin_div<T>(vry, vrr * sintht);
  }
  const T ux1 = uz * sintht + ux * costht, uy1 = uy, uz1 = uz * costht - ux * sintht;
  const T ux2 = ux1 * cosphi - uy1 * sinphi, uy2 = ux1 * sinphi + uy1 * cosphi, uz2 = uz1;
  const T vfx = vcx + vcr * ux2, vfy = vcy + vcr * uy2, vfz = vcz + vcr * uz2;
  const T vfr = kin_sqrt<T>(vfx * vfx + vfy * vfy + vfz * vfz);
  if (kin_div<T>(vfr, (T)pc.viT) > (T)10.0) return 2;
  vx = vfx; vy = vfy; vz = vfz;
  return 1;
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}


// Boris rotation (particles.c:374-447 with the intended cross products and
// per-component v-minus); operation order shared with the oracle.
template <typename T, typename PC>
__device__ __forceinline__ void boris(const PC &pc, T e0, T e1, T e2, T &vx, T &vy, T &vz)
{
  const T tau = (T)pc.tau, t0 = (T)pc.t[0], t1 = (T)pc.t[1], t2 = (T)pc.t[2];
  const T s0 = (T)pc.s[0], s1 = (T)pc.s[1], s2 = (T)pc.s[2];
  // fmad(a, b, -c) is a*b - c and fmad(tau, e, v) is v + tau*e exactly in FP64
  const T vm0 = fmad(tau, e0, vx), vm1 = fmad(tau, e1, vy), vm2 = fmad(tau, e2, vz);
  const T vp0 = vm0 + fmad(vm1, t2, -(vm2 * t1));
  const T vp1 = vm1 + fmad(vm2, t0, -(vm0 * t2));
  const T vp2 = vm2 + fmad(vm0, t1, -(vm1 * t0));
  const T vq0 = vm0 + fmad(vp1, s2, -(vp2 * s1));
  const T vq1 = vm1 + fmad(vp2, s0, -(vp0 * s2));
  const T vq2 = vm2 + fmad(vp0, s1, -(vp1 * s0));
  vx = fmad(tau, e0, vq0);
  vy = fmad(tau, e1, vq1);
  vz = fmad(tau, e2, vq2);
}

// Trilinear gather from the interleaved node field, corner order q = 0..7.
template <typename T>
__device__ __forceinline__ void gather(const Grid &g, const Vec4<T> *__restrict__ E4, int i, int j, int k, T fx, T fy,
                                       T fz, T &e0, T &e1, T &e2)
{
  T W[8];
  corner_weights<T>(fx, fy, fz, W);
  e0 = (T)0; e1 = (T)0; e2 = (T)0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const Vec4<T> e = E4[lin(g, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2))];
    e0 = fmad(W[q], e.x, e0);
    e1 = fmad(W[q], e.y, e1);
    e2 = fmad(W[q], e.z, e2);
  }
}

// The same gather for an ion of cell c = (i, j, k): corner indices from
// c and the periodic +1 offsets along each axis (no per-corner index math).
template <typename T>
__device__ __forceinline__ void gather_c(const Grid &g, const Vec4<T> *__restrict__ E4, int64_t c, int i, int j, int k,
                                         T fx, T fy, T fz, T &e0, T &e1, T &e2)
{
  T W[8];
  corner_weights<T>(fx, fy, fz, W);
  const int64_t ox = i + 1 == g.nx ? 1 - g.nx : 1;
  const int64_t oy = (j + 1 == g.ny ? 1 - g.ny : 1) * (int64_t)g.nx;
  const int64_t oz = (k + 1 == g.nz ? 1 - g.nz : 1) * (int64_t)g.nx * g.ny;
  e0 = (T)0; e1 = (T)0; e2 = (T)0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const Vec4<T> e = E4[c + ((q & 1) ? ox : 0) + ((q & 2) ? oy : 0) + ((q & 4) ? oz : 0)];
    e0 = fmad(W[q], e.x, e0);
    e1 = fmad(W[q], e.y, e1);
    e2 = fmad(W[q], e.z, e2);
  }
}

// Mixed move: in-cell offset += v dt / d with an exact integer carry, for an
// ion of cell (i, j, k); returns the new cell and leaves its indices in (i, j, k).
template <typename PC>
__device__ __forceinline__ int move_mixed_ijk(const Grid &g, const PC &pc, int &i, int &j, int &k, float &fx,
                                              float &fy, float &fz, float vx, float vy, float vz)
{
  float o[3] = {fmaf(vx, (float)pc.cdt[0], fx), fmaf(vy, (float)pc.cdt[1], fy), fmaf(vz, (float)pc.cdt[2], fz)};
  int ci[3] = {i, j, k};
  const int nn[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float fl = floorf(o[a]);
    float f = o[a] - fl;
    int sh = (int)fl;
    if (f >= 1.0f) { f -= 1.0f; sh += 1; }
    int cc = ci[a] + sh;
    if (cc < 0 || cc >= nn[a]) {  // more than one period away: full modulo
      cc %= nn[a];
      if (cc < 0) cc += nn[a];
    }
    ci[a] = cc;
    o[a] = f;
  }
  fx = o[0]; fy = o[1]; fz = o[2];
  i = ci[0]; j = ci[1]; k = ci[2];
  return (int)(ci[0] + (int64_t)g.nx * (ci[1] + (int64_t)g.ny * ci[2]));
}

// FP64 absolute move and cell key.
__device__ __forceinline__ int move_fp64(const Grid &g, double dt, double &x, double &y, double &z, double vx,
                                         double vy, double vz)
{
  x = wrap_pos(x + vx * dt, g.p0[0], g.p1[0], g.L[0]);
  y = wrap_pos(y + vy * dt, g.p0[1], g.p1[1], g.L[1]);
  z = wrap_pos(z + vz * dt, g.p0[2], g.p1[2], g.L[2]);
  int i, j, k;
  double f;
  cell_coord(x, g.p0[0], g.d[0], g.nx, i, f);
  cell_coord(y, g.p0[1], g.d[1], g.ny, j, f);
  cell_coord(z, g.p0[2], g.d[2], g.nz, k, f);
  return (int)(i + (int64_t)g.nx * (j + (int64_t)g.ny * k));
}

}  // namespace chy
