// chy_particles.hip -- particle kernels for gfx950.
//
//  push     gather of E = E0 - grad(phi) (trilinear, the transpose of the
//           deposit), Boris rotation, ion-neutral collisions and the periodic
//           move, in one pass over cell-sorted particles (reference
//           particles.c:347-718, vectors.c:12-162);
//  sort     stable counting sort by cell: integer histogram (fused into the
//           push), exclusive scan, wave-aggregated scatter, then a per-cell
//           rank by source index that restores the stable order, so the
//           result is deterministic although the scatter uses atomics
//           (replaces DMSwarmMigrate, particles.c:714);
//  deposit  CIC moments without floating-point atomics: per-cell corner
//           partial sums in sorted order, then per-node sums over the 8
//           adjacent cells in fixed corner order (particles.c:136-255).
//
// FP64 state: one array per field -- absolute x,y,z, v (f64), pid, optional
// weight, cell key.  Mixed state: per ion the 16-byte records
// (ox, oy, oz, w) and (vx, vy, vz, id) of FP32 in-cell offsets, weight,
// velocity and id bits, and the int32 cell key.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "chy_internal.h"
#include "chy_push.h"

namespace chy {

template <typename T> void launch_fill(chy_ctx *ctx, int64_t n, T *v, T val);

// Workgroups are dealt to the 8 XCDs round-robin (block b on XCD b % 8).  The
// grid-stride particle kernels index by this logical block instead, so that
// the blocks of one XCD cover one contiguous eighth of every grid-wide window
// and ions, their neighbours and their field corners share that XCD's L2.
__device__ __forceinline__ int64_t xcd_block()
{
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  if (nb & 7u) return b;
  return (int64_t)(b & 7u) * (nb >> 3) + (b >> 3);
}

// Step between two cell indices along an axis of n cells: 0, +1, -1 (with
// the periodic wrap), or 2 when they are not neighbours.
__device__ __forceinline__ int nb_step(int diff, int n)
{
  if (diff == 0) return 0;
  if (diff == 1 || diff == 1 - n) return 1;
  if (diff == -1 || diff == n - 1) return -1;
  return 2;  // not a neighbour
}

// Wave-aggregated integer atomics on per-cell counters for runs of equal keys.
__device__ __forceinline__ void count_key(uint32_t *counts, int key)
{
  int head, len;
  key_runs(key, head, len);
  if (key >= 0 && lane_id() == head) atomicAdd(&counts[key], (uint32_t)len);
}

__device__ __forceinline__ void count_flags(unsigned long long *cs, int slot, bool f)
{
  const uint64_t m = __ballot(f);
  if (m && lane_id() == 0) atomicAdd(&cs[slot], (unsigned long long)__popcll(m));
}

// The block's collision tallies -- accepted and refused scatterings, and the
// sums of the collision probabilities p and of p (1 - p) (the mean and
// variance of its Bernoulli count) -- to bpart[4 block + q], q = 0..3, for
// k_push_reduce to add in a fixed order: no atomics on shared addresses
// (one per block had cost a few ns each, serialised), and the FP64 sums are
// reproducible.
__device__ __forceinline__ void flush_block(double *bpart, uint32_t nacc, uint32_t nref, double pm, double pv,
                                            double *sh)
{
  double v[4] = {(double)nacc, (double)nref, pm, pv};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] += __shfl_down(v[q], off, 64);
  }
  if (lane_id() == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) sh[4 * (threadIdx.x >> 6) + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    bpart[4 * (int64_t)blockIdx.x + q] = (sh[q] + sh[4 + q]) + (sh[8 + q] + sh[12 + q]);
  }
}

// Sums of the push's block tallies in a fixed order, in two levels: FINAL
// false, block g (of kPushRed) adds blocks [g kPushRedSpan, (g+1)
// kPushRedSpan) of bpart into out[4 g + q]; FINAL true, one block adds the
// kPushRed results and writes cs[0] accepted, cs[1] refused (counts), cs[4]
// and cs[5] the FP64 sums of p and p (1 - p).
constexpr int kPushRed = 64, kPushRedSpan = kPushBlocks / kPushRed;
template <bool FINAL>
__global__ __launch_bounds__(256) void k_push_reduce(const double *__restrict__ in, int n, double *__restrict__ out,
                                                     unsigned long long *__restrict__ cs)
{
  __shared__ double sh[4][4];
  const int lo = FINAL ? 0 : (int)blockIdx.x * kPushRedSpan;
  const int hi = FINAL ? n : (lo + kPushRedSpan < n ? lo + kPushRedSpan : n);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = lo + (int)threadIdx.x; b < hi; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] += in[4 * (int64_t)b + q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] += __shfl_down(v[q], off, 64);
  }
  if (lane_id() == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) sh[threadIdx.x >> 6][q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    const double t = (sh[0][q] + sh[1][q]) + (sh[2][q] + sh[3][q]);
    if (!FINAL) out[4 * (int64_t)blockIdx.x + q] = t;
    else if (q < 2) cs[q] = (unsigned long long)(t + 0.5);
    else reinterpret_cast<double *>(cs)[2 + q] = t;
  }
}

// FP64 push: the block's collision moments and scattering counts by global
// atomics (one per block and value).  (Block partials, as the FP32 push
// writes them, took this kernel from 11.7 to 12.3 ms: it is held to three
// waves per SIMD with a few spills, and its schedule shifted.)
__device__ __forceinline__ void flush_collision_moments(double *cexp, double pm, double pv, double *sh)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    pm += __shfl_down(pm, off, 64);
    pv += __shfl_down(pv, off, 64);
  }
  if (lane_id() == 0) { sh[2 * (threadIdx.x >> 6)] = pm; sh[2 * (threadIdx.x >> 6) + 1] = pv; }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int q = threadIdx.x;
    atomicAdd(&cexp[q], (sh[q] + sh[2 + q]) + (sh[4 + q] + sh[6 + q]));
  }
}

__device__ __forceinline__ void flush_scatter_counts(unsigned long long *cs, uint32_t nacc, uint32_t nref,
                                                     unsigned int *sacc)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    nacc += __shfl_down(nacc, off, 64);
    nref += __shfl_down(nref, off, 64);
  }
  if (lane_id() == 0) {
    if (nacc) atomicAdd(&sacc[0], nacc);
    if (nref) atomicAdd(&sacc[1], nref);
  }
  __syncthreads();
  if (threadIdx.x < 2 && sacc[threadIdx.x]) atomicAdd(&cs[threadIdx.x], (unsigned long long)sacc[threadIdx.x]);
}

// ---- main push: gather + Boris + collision (decision and scattering) + move -----------
// One neutral draw per ion serves both the decision and, for the ions that
// collide, the scattering, which runs in place: every ion's new state is
// written by the same store instructions as its neighbours' (no scattered
// write-back of the colliders).
// At least 3 waves per SIMD: 168 registers (a few spills) against 178 at 2.
template <bool COLLIDE, bool MOVE>
__global__ __launch_bounds__(256, 3) void k_push_fp64(Grid g, PushConst pc, int64_t np, PState<double> st,
                                                   const Vec4<double> *__restrict__ E4, const double *__restrict__ Sdev,
                                                   uint32_t *__restrict__ counts, unsigned long long *__restrict__ cstats,
                                                   double *__restrict__ bpart, uint8_t *__restrict__ dir)
{
  __shared__ unsigned int sacc[2];
  __shared__ double smom[8];
  double pm = 0.0, pv = 0.0;  // sums of the collision probabilities p and of p (1 - p)
  if (threadIdx.x < 2) sacc[threadIdx.x] = 0u;
  __syncthreads();
  const double S = COLLIDE ? Sdev[0] : 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // every lane iterates the same number of times so wave-level collectives stay convergent
  const int64_t iters = (np + stride - 1) / stride;
  uint32_t nacc = 0, nref = 0;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + xcd_block() * blockDim.x + threadIdx.x;
    const bool live = p < np;
    int key = -1, capped = 0;
    bool bad = false;
    if (live) {
      double x = st.x[p], y = st.y[p], z = st.z[p];
      double vx = st.vx[p], vy = st.vy[p], vz = st.vz[p];
      int i, j, k;
      double fx, fy, fz, e0, e1, e2;
      cell_coord(x, g.p0[0], g.d[0], g.nx, i, fx);
      cell_coord(y, g.p0[1], g.d[1], g.ny, j, fy);
      cell_coord(z, g.p0[2], g.d[2], g.nz, k, fz);
      gather_c<double>(g, E4, (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * k), i, j, k, fx, fy, fz, e0,
                       e1, e2);
      boris<double>(pc, e0, e1, e2, vx, vy, vz);
      if (COLLIDE) {
        const uint32_t pid = st.pid[p];
        double vnx, vny, vnz, ud;
        neutral_draw<double>(pc, pid, vnx, vny, vnz, ud);
        const double pr = collide_prob_n<double>(pc, S, vnx, vny, vnz, vx, vy, vz, capped);
        pm += pr;
        pv += pr * (1.0 - pr);
        if (ud < pr) {
          if (scatter_n<double>(pc, pid, vnx, vny, vnz, vx, vy, vz) == 1) nacc++;
          else nref++;
        }
      }
      st.vx[p] = vx; st.vy[p] = vy; st.vz[p] = vz;
      if (MOVE) {
        key = move_fp64(g, pc.dt, x, y, z, vx, vy, vz);
        if (dir) {  // the move direction for the neighbour-move sort (k_trans_dir)
          int i2, j2, k2;
          split_cell(g, key, i2, j2, k2);
          const int di = nb_step(i2 - i, g.nx), dj = nb_step(j2 - j, g.ny), dk = nb_step(k2 - k, g.nz);
          dir[p] = (di == 2 || dj == 2 || dk == 2) ? (uint8_t)255
                                                   : (uint8_t)((di + 1) + 3 * (dj + 1) + 9 * (dk + 1));
        }
        st.x[p] = x; st.y[p] = y; st.z[p] = z;
        st.cell[p] = key;
      }
      bad = !isfinite(vx) || !isfinite(vy) || !isfinite(vz) || !isfinite(x) || !isfinite(y) || !isfinite(z);
    }
    if (MOVE && counts) count_key(counts, key);
    if (COLLIDE) count_flags(cstats, 2, capped != 0);
    count_flags(cstats, 3, bad);
  }
  if (COLLIDE) {
    flush_scatter_counts(cstats, nacc, nref, sacc);
    flush_collision_moments(reinterpret_cast<double *>(cstats + 4), pm, pv, smom);
  }
}

template <bool COLLIDE, bool MOVE>
__global__ __launch_bounds__(256) void k_push_mixed(Grid g, PushConstF pc, int64_t np, PState<float> st,
                                                    const Vec4<float> *__restrict__ E4, const double *__restrict__ Sdev,
                                                    uint32_t *__restrict__ counts, unsigned long long *__restrict__ cstats,
                                                    double *__restrict__ bpart, uint8_t *__restrict__ dir)
{
  __shared__ double smom[16];
  float pm = 0.f, pv = 0.f;  // this thread's sums of p and p (1 - p) (a few ions each)
  const float S = COLLIDE ? (float)Sdev[0] : 0.f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t iters = (np + stride - 1) / stride;
  uint32_t nacc = 0, nref = 0;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + xcd_block() * blockDim.x + threadIdx.x;
    const bool live = p < np;
    int key = -1, capped = 0;
    bool bad = false;
    if (live) {
      const int c = st.cell[p];
      int i, j, k;
      split_cell(g, c, i, j, k);
      const Vec4<float> P = st.pos[p], V = st.vel[p];
      const float fx = P.x, fy = P.y, fz = P.z;
      float vx = V.x, vy = V.y, vz = V.z;
      float e0, e1, e2;
      gather_c<float>(g, E4, c, i, j, k, fx, fy, fz, e0, e1, e2);
      boris<float>(pc, e0, e1, e2, vx, vy, vz);
      if (COLLIDE) {
        const uint32_t pid = __float_as_uint(V.w);
        float vnx, vny, vnz, ud;
        neutral_draw<float>(pc, pid, vnx, vny, vnz, ud);
        const float pr = collide_prob_n<float>(pc, S, vnx, vny, vnz, vx, vy, vz, capped);
        pm += pr;
        pv += pr * (1.0f - pr);
        if (ud < pr) {
          if (scatter_n<float>(pc, pid, vnx, vny, vnz, vx, vy, vz) == 1) nacc++;
          else nref++;
        }
      }
      st.vel[p] = Vec4<float>{vx, vy, vz, V.w};
      if (MOVE) {
        float ox = fx, oy = fy, oz = fz;
        const int i0 = i, j0 = j, k0 = k;
        key = move_mixed_ijk(g, pc, i, j, k, ox, oy, oz, vx, vy, vz);
        if (dir) {  // the move direction for the neighbour-move sort (k_trans_dir)
          const int di = nb_step(i - i0, g.nx), dj = nb_step(j - j0, g.ny), dk = nb_step(k - k0, g.nz);
          dir[p] = (di == 2 || dj == 2 || dk == 2) ? (uint8_t)255
                                                   : (uint8_t)((di + 1) + 3 * (dj + 1) + 9 * (dk + 1));
        }
        st.pos[p] = Vec4<float>{ox, oy, oz, P.w};
        st.cell[p] = key;
        bad = !isfinite(ox) || !isfinite(oy) || !isfinite(oz);
      }
      bad = bad || !isfinite(vx) || !isfinite(vy) || !isfinite(vz) || !isfinite(fx) || !isfinite(fy) || !isfinite(fz);
    }
    if (MOVE && counts) count_key(counts, key);
    if (COLLIDE) count_flags(cstats, 2, capped != 0);
    count_flags(cstats, 3, bad);
  }
  if (COLLIDE) {
    flush_block(bpart, nacc, nref, (double)pm, (double)pv, smom);
  }
}

// Move only (positions from the current velocities), with keys and counts.
__global__ __launch_bounds__(256) void k_move_fp64(Grid g, PushConst pc, int64_t np, PState<double> st,
                                                   uint32_t *__restrict__ counts, unsigned long long *__restrict__ cstats)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t iters = (np + stride - 1) / stride;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    bool bad = false;
    if (p < np) {
      double x = st.x[p], y = st.y[p], z = st.z[p];
      key = move_fp64(g, pc.dt, x, y, z, st.vx[p], st.vy[p], st.vz[p]);
      st.x[p] = x; st.y[p] = y; st.z[p] = z;
      st.cell[p] = key;
      bad = !isfinite(x) || !isfinite(y) || !isfinite(z);
    }
    count_key(counts, key);
    count_flags(cstats, 3, bad);
  }
}

__global__ __launch_bounds__(256) void k_move_mixed(Grid g, PushConstF pc, int64_t np, PState<float> st,
                                                    uint32_t *__restrict__ counts, unsigned long long *__restrict__ cstats)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t iters = (np + stride - 1) / stride;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    bool bad = false;
    if (p < np) {
      const Vec4<float> P = st.pos[p], V = st.vel[p];
      float fx = P.x, fy = P.y, fz = P.z;
      int i, j, k;
      split_cell(g, st.cell[p], i, j, k);
      key = move_mixed_ijk(g, pc, i, j, k, fx, fy, fz, V.x, V.y, V.z);
      st.pos[p] = Vec4<float>{fx, fy, fz, P.w};
      st.cell[p] = key;
      bad = !isfinite(fx) || !isfinite(fy) || !isfinite(fz);
    }
    count_key(counts, key);
    count_flags(cstats, 3, bad);
  }
}

// Collision normaliser partials over the current velocities.
template <typename T>
__global__ __launch_bounds__(256) void k_meanrel(int64_t np, const T *__restrict__ vx, const T *__restrict__ vy,
                                                 const T *__restrict__ vz, double n0x, double n0y, double n0z,
                                                 double vnT, double *__restrict__ sred)
{
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
    acc += mean_rel_speed((double)vx[p], (double)vy[p], (double)vz[p], n0x, n0y, n0z, vnT);
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) sred[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_meanrel_mixed(int64_t np, const Vec4<float> *__restrict__ vel, double n0x,
                                                       double n0y, double n0z, double vnT, double *__restrict__ sred)
{
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
    const Vec4<float> v = vel[p];
    acc += mean_rel_speed((double)v.x, (double)v.y, (double)v.z, n0x, n0y, n0z, vnT);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) sred[blockIdx.x] = s;
}

// FP64 keys and histogram from absolute positions.
__global__ __launch_bounds__(256) void k_keys_fp64(Grid g, int64_t np, PState<double> st, uint32_t *__restrict__ counts)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t iters = (np + stride - 1) / stride;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    if (p < np) {
      int i, j, k;
      double f;
      cell_coord(st.x[p], g.p0[0], g.d[0], g.nx, i, f);
      cell_coord(st.y[p], g.p0[1], g.d[1], g.ny, j, f);
      cell_coord(st.z[p], g.p0[2], g.d[2], g.nz, k, f);
      key = (int)(i + (int64_t)g.nx * (j + (int64_t)g.ny * k));
      st.cell[p] = key;
    }
    count_key(counts, key);
  }
}

// Histogram from stored keys (mixed state).
__global__ __launch_bounds__(256) void k_keys_count(int64_t np, const int32_t *__restrict__ cell, uint32_t *__restrict__ counts)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t iters = (np + stride - 1) / stride;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    count_key(counts, p < np ? cell[p] : -1);
  }
}

// ---- exclusive scan of the cell histogram ------------------------------------------
constexpr int kScanTile = 4096;  // 256 threads x 16
constexpr uint32_t kMaxCellOccupancy = 8192;

// Tile sums and tile maxima of the cell counts (no atomics: the single-block
// scan below takes the largest of the tile maxima).
__global__ __launch_bounds__(256) void k_tile_sum(int64_t n, const uint32_t *__restrict__ in, uint32_t *__restrict__ tsum,
                                                  uint32_t *__restrict__ tmax)
{
  __shared__ uint32_t sh[4], shm[4];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t v[kScanTile / 256];
#pragma unroll
  for (int u = 0; u < kScanTile / 256; u++) {
    const int64_t i = base + threadIdx.x + 256 * u;
    v[u] = i < n ? in[i] : 0u;
  }
  uint32_t acc = 0, mx = 0;
#pragma unroll
  for (int u = 0; u < kScanTile / 256; u++) { acc += v[u]; mx = v[u] > mx ? v[u] : mx; }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_down(mx, off, 64);
    mx = o > mx ? o : mx;
    acc += __shfl_down(acc, off, 64);
  }
  if ((threadIdx.x & 63) == 0) { sh[threadIdx.x >> 6] = acc; shm[threadIdx.x >> 6] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    tsum[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    const uint32_t m01 = shm[0] > shm[1] ? shm[0] : shm[1], m23 = shm[2] > shm[3] ? shm[2] : shm[3];
    tmax[blockIdx.x] = m01 > m23 ? m01 : m23;
  }
}

// Block-wide exclusive scan of 256 values (one per thread); returns the total.
__device__ __forceinline__ uint32_t block_exscan(uint32_t v, uint32_t &total, uint32_t *sh)
{
  const int l = lane_id(), w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (l >= off) x += y;
  }
  if (l == 63) sh[w] = x;
  __syncthreads();
  uint32_t wbase = 0;
  for (int q = 0; q < w; q++) wbase += sh[q];
  total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return wbase + x - v;
}

// Single block: exclusive scan of the tile sums in place, and the largest
// tile maximum (the largest cell occupancy) into maxcount[0].
__global__ __launch_bounds__(256) void k_scan_tiles(int64_t nt, uint32_t *__restrict__ tsum,
                                                    const uint32_t *__restrict__ tmax, uint32_t *__restrict__ maxcount)
{
  __shared__ uint32_t sh[4];
  uint32_t mx = 0;
  for (int64_t i = threadIdx.x; i < nt; i += 256) mx = tmax[i] > mx ? tmax[i] : mx;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_down(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t m01 = sh[0] > sh[1] ? sh[0] : sh[1], m23 = sh[2] > sh[3] ? sh[2] : sh[3];
    maxcount[0] = m01 > m23 ? m01 : m23;
  }
  __syncthreads();
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nt; b0 += 256) {
    const int64_t i = b0 + threadIdx.x;
    const uint32_t v = i < nt ? tsum[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exscan(v, tot, sh);
    if (i < nt) tsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) tsum[nt] = carry;  // the histogram's total, checked against the ion count
}

// Per-tile scan; writes offsets[0..n) and the cursor copy; offsets[n] = total.
__global__ __launch_bounds__(256) void k_tile_scan(int64_t n, const uint32_t *__restrict__ in,
                                                   const uint32_t *__restrict__ tsum, uint32_t *__restrict__ offs,
                                                   uint32_t *__restrict__ cursor, uint32_t total)
{
  __shared__ uint32_t sh[4];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t carry = tsum[blockIdx.x];
  for (int q0 = 0; q0 < kScanTile; q0 += 256) {
    const int64_t i = base + q0 + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exscan(v, tot, sh);
    if (i < n) { offs[i] = carry + ex; cursor[i] = carry + ex; }
    carry += tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) offs[n] = total;
}

// Scatter: slot = cursor[key]++ (run-aggregated), perm[slot] = p.
__global__ __launch_bounds__(256) void k_scatter(int64_t np, const int32_t *__restrict__ cell, uint32_t *__restrict__ cursor,
                                                 uint32_t *__restrict__ perm, int32_t *__restrict__ slotkey,
                                                 uint32_t *__restrict__ lost)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t iters = (np + stride - 1) / stride;
  for (int64_t it = 0; it < iters; it++) {
    const int64_t p = it * stride + xcd_block() * blockDim.x + threadIdx.x;
    const int key = p < np ? cell[p] : -1;
    int head, len;
    key_runs(key, head, len);
    uint32_t base = 0;
    if (key >= 0 && lane_id() == head) base = atomicAdd(&cursor[key], (uint32_t)len);
    base = __shfl(base, head, 64);
    const uint32_t slot = base + (uint32_t)(lane_id() - head);
    if (key >= 0 && slot < np) {  // (the host has checked the histogram's total)
      perm[slot] = (uint32_t)p;
      slotkey[slot] = key;
    } else if (key >= 0) {
      atomicOr(lost, 1u);  // a slot outside the ion range: reported by the next sort
    }
  }
}

// Stable order within each cell: rank of a slot's source index among its
// cell's sources; moves the particle from generation `a` to `b`.  With SUMS
// it also accumulates the collision normaliser over the final velocities in
// this deterministic (sorted) order, one partial per block.
template <typename T, bool SUMS>
__global__ __launch_bounds__(256) void k_reorder(int64_t np, const uint32_t *__restrict__ perm,
                                                 const int32_t *__restrict__ slotkey, const uint32_t *__restrict__ offs,
                                                 PState<T> a, PState<T> b, bool weighted, double n0x, double n0y,
                                                 double n0z, double vnT, double *__restrict__ sred,
                                                 uint32_t *__restrict__ src)
{
  // a block takes 256-slot chunks; the permutation of the chunk and of
  // kRankHalo slots either side is staged in LDS for the per-cell ranks
  constexpr int kRankHalo = 128, kWin = 256 + 2 * kRankHalo;
  __shared__ uint32_t pw[kWin];
  __shared__ double sh[4];
  double acc = 0.0;
  for (int64_t s0 = xcd_block() * 256; s0 < np; s0 += (int64_t)gridDim.x * 256) {
    const int64_t wlo = s0 - kRankHalo;
    for (int e = threadIdx.x; e < kWin; e += blockDim.x) {
      const int64_t t = wlo + e;
      pw[e] = (t >= 0 && t < np) ? perm[t] : 0u;
    }
    __syncthreads();
    const int64_t s = s0 + threadIdx.x;
    if (s < np) {
      const int c = slotkey[s];
      const uint32_t lo = offs[c], hi = offs[c + 1];
      const uint32_t me = pw[s - wlo];
      uint32_t rank = 0;
      if ((int64_t)lo >= wlo && (int64_t)hi <= wlo + kWin) {
        const uint32_t *q = pw + (lo - wlo);
        const int len = (int)(hi - lo);
        uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        int t = 0;
        for (; t + 4 <= len; t += 4) {
          r0 += q[t] < me; r1 += q[t + 1] < me; r2 += q[t + 2] < me; r3 += q[t + 3] < me;
        }
        for (; t < len; t++) r0 += q[t] < me;
        rank = (r0 + r1) + (r2 + r3);
      } else {
        for (uint32_t t = lo; t < hi; t++) rank += perm[t] < me ? 1u : 0u;
      }
      const uint32_t d = lo + rank;
      if (src) src[d] = me;  // source index of each slot, when asked for
      T vx, vy, vz;
      if constexpr (sizeof(T) == 4) {
        const Vec4<float> P = a.pos[me], V = a.vel[me];
        vx = V.x; vy = V.y; vz = V.z;
        b.pos[d] = P; b.vel[d] = V;
      } else {
        vx = a.vx[me]; vy = a.vy[me]; vz = a.vz[me];
        b.x[d] = a.x[me]; b.y[d] = a.y[me]; b.z[d] = a.z[me];
        b.vx[d] = vx; b.vy[d] = vy; b.vz[d] = vz;
        b.pid[d] = a.pid[me];
        if (weighted) b.w[d] = a.w[me];
      }
      b.cell[d] = c;
      if (SUMS) {
        if (sizeof(T) == 4) acc += (double)mean_rel_speed_f((float)vx, (float)vy, (float)vz, (float)n0x, (float)n0y,
                                                           (float)n0z, (float)vnT);
        else acc += mean_rel_speed((double)vx, (double)vy, (double)vz, n0x, n0y, n0z, vnT);
      }
    }
    __syncthreads();
  }
  if (SUMS) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) sred[blockIdx.x] = t;
  }
}

// ---- neighbour-move sort ----------------------------------------------------------------
// After a push from the stable cell order every ion is in its old cell or in
// one of the 26 neighbours.  Its slot in the new stable order (cell, then old
// index) is then
//   offs[c] + (ions entering c from old cells numbered below its old cell,
//              the stayers of c included when c is below)
//           + (its rank among the ions of its old cell that go to c).
// From the new keys, k_trans counts, per old cell, the ions going each way and
// ranks every ion within its direction; k_counts turns the counts into the new histogram and
// the start of every (old cell, direction) group within its new cell; k_perm
// places every ion (perm[slot] = old index) and k_cell_partials<GATHER>
// moves the state in slot order and makes the deposit's cell partials from it.
// The atomic scatter and the per-cell ranking of the general sort are not
// needed.  A move further than one cell, or 255 ions of one cell going one way,
// sets a flag and the general sort runs instead.  When the push wrote the move
// directions, k_trans_dir (below) reads them in place of the keys.


__device__ __forceinline__ int64_t nb_cell(const Grid &g, int i, int j, int k)
{
  return (int64_t)wrapi(i, g.nx) + (int64_t)g.nx * ((int64_t)wrapi(j, g.ny) + (int64_t)g.ny * wrapi(k, g.nz));
}

// direction index d = (di + 1) + 3 (dj + 1) + 9 (dk + 1); 13 = stay
__global__ __launch_bounds__(256) void k_trans(Grid g, const uint32_t *__restrict__ oldoffs,
                                               const int32_t *__restrict__ cell, uint16_t *__restrict__ info,
                                               uint8_t *__restrict__ trans, uint32_t *__restrict__ flag)
{
  __shared__ uint16_t cnt[27][256];  // column t belongs to thread t
  const int t = threadIdx.x;
  for (int d = 0; d < 27; d++) cnt[d][t] = 0;
  const int64_t c = xcd_block() * 256 + t;
  if (c >= g.nc) return;
  int i, j, k;
  split_cell(g, (int)c, i, j, k);
  bool bad = false;
  const uint32_t lo = oldoffs[c], hi = oldoffs[c + 1];
  constexpr int kB = 8;  // keys in flight per thread
  for (uint32_t q0 = lo; q0 < hi; q0 += kB) {
    int key[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) key[u] = q0 + u < hi ? cell[q0 + u] : 0;
#pragma unroll
    for (int u = 0; u < kB; u++) {
      if (q0 + u >= hi) break;
      int i2, j2, k2;
      split_cell(g, key[u], i2, j2, k2);
      const int di = nb_step(i2 - i, g.nx), dj = nb_step(j2 - j, g.ny), dk = nb_step(k2 - k, g.nz);
      if (di == 2 || dj == 2 || dk == 2) { bad = true; continue; }
      const int d = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);
      const uint32_t r = cnt[d][t];
      if (r >= 255) { bad = true; continue; }
      info[q0 + u] = (uint16_t)((d << 8) | r);
      cnt[d][t] = (uint16_t)(r + 1);
    }
  }
  for (int d = 0; d < 27; d++) trans[(int64_t)d * g.nc + c] = (uint8_t)cnt[d][t];
  if (bad) atomicOr(flag, 1u);
}

// The three source coordinates of new coordinate c along an axis of n cells,
// in ascending wrapped order, as moves m (source = c - m, m in {-1, 0, 1}).
// Cell numbers are ordered z, then y, then x, so nesting the three axes'
// orders lists the 27 sources of a cell in ascending cell number.
__device__ __forceinline__ void src_order(int c, int n, int m[3])
{
  if (c == 0 && n > 1) { m[0] = 0; m[1] = -1; m[2] = 1; }        // c, c + 1, then c - 1 = n - 1
  else if (c == n - 1 && n > 1) { m[0] = -1; m[1] = 1; m[2] = 0; } // c + 1 = 0, c - 1, c
  else { m[0] = 1; m[1] = 0; m[2] = -1; }                       // c - 1, c, c + 1
}

// counts[c] = ions ending in c; base[d][s] = the ions entering c = s + off(d)
// from old cells numbered below s (the group of ions going from s in
// direction d starts there within c's segment): a running sum over the 27
// sources in ascending cell number, their 27 group sizes loaded first.
__global__ __launch_bounds__(256) void k_counts(Grid g, const uint8_t *__restrict__ trans, uint32_t *__restrict__ counts,
                                                uint16_t *__restrict__ base)
{
  const int64_t c = xcd_block() * 256 + threadIdx.x;
  if (c >= g.nc) return;
  int i, j, k;
  split_cell(g, (int)c, i, j, k);
  int mx[3], my[3], mz[3];
  src_order(i, g.nx, mx);
  src_order(j, g.ny, my);
  src_order(k, g.nz, mz);
  uint32_t t[27];
#pragma unroll
  for (int p = 0; p < 27; p++) {
    const int di = mx[p % 3], dj = my[(p / 3) % 3], dk = mz[p / 9];
    const int d = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);
    t[p] = trans[(int64_t)d * g.nc + nb_cell(g, i - di, j - dj, k - dk)];
  }
  uint32_t acc = 0;
#pragma unroll
  for (int p = 0; p < 27; p++) {
    const int di = mx[p % 3], dj = my[(p / 3) % 3], dk = mz[p / 9];
    const int d = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);
    base[(int64_t)d * g.nc + nb_cell(g, i - di, j - dj, k - dk)] = (uint16_t)acc;
    acc += t[p];
  }
  counts[c] = acc;
}

__global__ __launch_bounds__(256) void k_perm(Grid g, int64_t np, const int32_t *__restrict__ cell,
                                              const uint16_t *__restrict__ info, const uint16_t *__restrict__ base,
                                              const uint32_t *__restrict__ offs, uint32_t *__restrict__ perm,
                                              uint32_t *__restrict__ lost)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = xcd_block() * blockDim.x + threadIdx.x; p < np; p += stride) {
    const int c = cell[p];
    const uint32_t inf = info[p];
    const int d = (int)(inf >> 8);
    int i, j, k;
    split_cell(g, c, i, j, k);
    const int64_t co = nb_cell(g, i - (d % 3 - 1), j - ((d / 3) % 3 - 1), k - (d / 9 - 1));  // old cell
    const uint32_t slot = offs[c] + base[(int64_t)d * g.nc + co] + (inf & 255u);
    if (slot < np) perm[slot] = (uint32_t)p;
    else atomicOr(lost, 1u);  // a slot outside the ion range: reported by the next sort
  }
}


// The directions of a block's old cells (one contiguous range, ions s0 ..
// s1) staged in LDS with aligned 4-byte loads: word w0 = s0 / 4 first, so
// ion q's byte is at q - 4 w0.  Returns false (nothing staged) when the
// range does not fit NW words per thread.
template <int NW>
__device__ __forceinline__ bool stage_dirs(const uint8_t *__restrict__ dir, uint32_t s0, uint32_t s1,
                                           uint32_t *__restrict__ sw)
{
  const uint32_t w0 = s0 >> 2, w1 = (s1 + 3) >> 2;
  if (w1 - w0 > (uint32_t)(NW * 256)) return false;
  const uint32_t *dw = reinterpret_cast<const uint32_t *>(dir);
  uint32_t v[NW];
#pragma unroll
  for (int u = 0; u < NW; u++) {
    const uint32_t w = w0 + threadIdx.x + 256u * u;
    v[u] = w < w1 ? dw[w] : 0u;
  }
#pragma unroll
  for (int u = 0; u < NW; u++) {
    const uint32_t w = w0 + threadIdx.x + 256u * u;
    if (w < w1) sw[w - w0] = v[u];
  }
  return true;
}

// k_trans from the move directions the push wrote (dir, one byte per ion in
// the old order): per old cell the 27 group sizes, and per ion its
// (direction, rank) record for k_perm.  The block's 256 cells' directions
// (one contiguous range) are staged in LDS by aligned coalesced loads, each
// thread walks its cell there, and the records go out coalesced from LDS.
// NW: staged words per thread; a block whose ions do not fit reads and
// writes global memory directly.
template <int NW>
__global__ __launch_bounds__(256) void k_trans_dir(Grid g, const uint32_t *__restrict__ oldoffs,
                                                   const uint8_t *__restrict__ dir, uint16_t *__restrict__ info,
                                                   uint8_t *__restrict__ trans, uint32_t *__restrict__ flag)
{
  constexpr int CAP = NW * 1024 - 8;  // ions whose directions always fit, whatever their alignment
  __shared__ uint32_t sw[NW * 256];
  __shared__ uint16_t sinfo[CAP];
  __shared__ uint16_t cnt[27][256];  // column t belongs to thread t
  __shared__ uint32_t so[257];
  const int t = (int)threadIdx.x;
  const int64_t c0 = xcd_block() * 256;
  if (c0 >= g.nc) return;
  const int ncell = (int)((g.nc - c0) < 256 ? (g.nc - c0) : 256);
  if (t <= ncell) so[t] = oldoffs[c0 + t];
  if (t == 0 && ncell == 256) so[256] = oldoffs[c0 + 256];
  for (int d = 0; d < 27; d++) cnt[d][t] = 0;
  __syncthreads();
  const uint32_t s0 = so[0], s1 = so[ncell];
  const bool staged = s1 - s0 <= (uint32_t)CAP && stage_dirs<NW>(dir, s0, s1, sw);
  __syncthreads();
  const uint8_t *sb = reinterpret_cast<const uint8_t *>(sw) + (s0 - 4u * (s0 >> 2));  // ion s0's direction
  bool bad = false;
  if (t < ncell) {
    const uint32_t lo = so[t], hi = so[t + 1];
    for (uint32_t q = lo; q < hi; q++) {
      const uint32_t d = staged ? (uint32_t)sb[q - s0] : (uint32_t)dir[q];
      if (d > 26u) { bad = true; continue; }
      const uint32_t r = cnt[d][t];
      if (r >= 255u) { bad = true; continue; }
      const uint16_t rec = (uint16_t)((d << 8) | r);
      if (staged) sinfo[q - s0] = rec;
      else info[q] = rec;
      cnt[d][t] = (uint16_t)(r + 1);
    }
    for (int d = 0; d < 27; d++) trans[(int64_t)d * g.nc + c0 + t] = (uint8_t)cnt[d][t];
  }
  if (staged) {
    __syncthreads();
    for (uint32_t e = (uint32_t)t; e < s1 - s0; e += 256) info[s0 + e] = sinfo[e];
  }
  if (bad) atomicOr(flag, 1u);
}

// ---- deposit ---------------------------------------------------------------------
// Per-cell corner partials P[(4q+m)*nc + c], m in {w, w vx, w vy, w vz}.
// A block owns kDepCells consecutive cells, whose particles are contiguous
// after the sort; they are staged through LDS in coalesced chunks of up to
// kDepUnroll * 256 ions, every load of a chunk issued before the first is
// used.  Eight threads share a cell, each summing one corner over the cell's
// particles in sorted order (the oracle's order), so the result is
// deterministic.
// GATHER: the same kernel finishes the neighbour-move sort -- slot s of the
// new order takes ion perm[s] of `st`, whose state is written to `b` (and
// src[s] = perm[s]), and the ion is deposited from LDS, so the sorted state
// is not read back; with SUMS the block's part of the collision normaliser
// over the final velocities in slot order goes to sred[block].
constexpr int kDepCells = 32;   // about 512 ions at 16 per cell: one chunk
constexpr int kDepUnroll = 3;   // staging passes of the block per chunk
constexpr int kDepChunk = 640;  // ions per chunk (LDS records)

// One staged ion: in-cell offsets, velocity and weight.
template <typename T> struct alignas(8 * sizeof(T)) DepRec { T fx, fy, fz, vx, vy, vz, w, pad; };

// XC (nx a multiple of kDepCells, so a block's cells lie in one x row): the
// block adds each cell's x-low corners to the x-high corners of the cell
// before it in LDS and writes 16 values per cell, P[(4r+m)*nc + c] for the
// y-z corner r = 2 (z bit) + (y bit), and the x-high corners of its last
// cell to E[(4r+m)*nblocks + block] (E = P + 16 nc), for the node sum to add
// at the block boundaries: half the partials' bytes written and read.
template <typename T, bool ABS, bool GATHER, bool SUMS, bool XC = false>
__global__ __launch_bounds__(256) void k_cell_partials(Grid g, const uint32_t *__restrict__ offs, PState<T> st,
                                                       bool weighted, T *__restrict__ P,
                                                       const uint32_t *__restrict__ perm, PState<T> b, double n0x,
                                                       double n0y, double n0z, double vnT, double *__restrict__ sred,
                                                       uint32_t *__restrict__ src)
{
  constexpr int kChunk = kDepChunk;
  static_assert(kChunk <= kDepUnroll * 256, "a chunk is staged in kDepUnroll passes of the block");
  __shared__ DepRec<T> rec[kChunk];
  __shared__ uint32_t so[kDepCells + 1];
  __shared__ double sh[4];
  double sacc = 0.0;
  const int64_t c0 = xcd_block() * kDepCells;
  const int ncell = (int)((g.nc - c0) < kDepCells ? (g.nc - c0) : kDepCells);
  if (threadIdx.x <= ncell) so[threadIdx.x] = offs[c0 + threadIdx.x];
  __syncthreads();
  const uint32_t s0 = so[0], s1 = so[ncell];
  const int cl = threadIdx.x >> 3, corner = threadIdx.x & 7;
  const bool own = cl < ncell;
  const uint32_t lo = own ? so[cl] : 0, hi = own ? so[cl + 1] : 0;
  T acc[4] = {(T)0, (T)0, (T)0, (T)0};
  for (uint32_t base = s0; base < s1; base += kChunk) {
    const int m = (int)((s1 - base) < (uint32_t)kChunk ? (s1 - base) : (uint32_t)kChunk);
    // A wave stages the passes u with a slot of the chunk (wave-uniform
    // count); each variant issues its indices first, then every field load,
    // then the writes, in straight-line code so all loads are in flight
    // together.  Slots past the chunk load ion `base` again (always present)
    // and write nothing.
    // the new cell of slot s is the block's cell whose segment holds s (the
    // sort put the ion there), so it is not read from the source
    auto cell_of = [&](uint32_t s) {
      int lo = 0, hi = ncell;  // so[lo] <= s < so[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (so[mid] <= s) lo = mid;
        else hi = mid;
      }
      return (int32_t)(c0 + lo);
    };
    auto stage = [&](auto passes) {
      constexpr int NU = decltype(passes)::value;
      uint32_t q[NU];
#pragma unroll
      for (int u = 0; u < NU; u++) {
        const int e = (int)threadIdx.x + 256 * u;
        const uint32_t s = base + (uint32_t)(e < m ? e : 0);
        q[u] = GATHER ? perm[s] : s;
      }
      if constexpr (sizeof(T) == 4) {
        // mixed: two 16-byte records and the key per ion
        Vec4<float> P[NU], V[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
          P[u] = st.pos[q[u]]; V[u] = st.vel[q[u]];
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
          const int e = (int)threadIdx.x + 256 * u;
          if (e < m) {
            const uint32_t s = base + (uint32_t)e;
            if (GATHER) {
              b.pos[s] = P[u]; b.vel[s] = V[u];
              b.cell[s] = cell_of(s);
              if (src) src[s] = q[u];
              if (SUMS)
                sacc += (double)mean_rel_speed_f(V[u].x, V[u].y, V[u].z, (float)n0x, (float)n0y, (float)n0z,
                                                 (float)vnT);
            }
            rec[e] = DepRec<T>{P[u].x, P[u].y, P[u].z, V[u].x, V[u].y, V[u].z, P[u].w, 0.0f};
          }
        }
      } else {
        T x[NU], y[NU], z[NU], vx[NU], vy[NU], vz[NU], w[NU];
        uint32_t pid[NU];
#pragma unroll
        for (int u = 0; u < NU; u++) {
          x[u] = st.x[q[u]]; y[u] = st.y[q[u]]; z[u] = st.z[q[u]];
          vx[u] = st.vx[q[u]]; vy[u] = st.vy[q[u]]; vz[u] = st.vz[q[u]];
          w[u] = weighted ? st.w[q[u]] : (T)1;
          if (GATHER) pid[u] = st.pid[q[u]];
        }
#pragma unroll
        for (int u = 0; u < NU; u++) {
          const int e = (int)threadIdx.x + 256 * u;
          if (e < m) {
            const uint32_t s = base + (uint32_t)e;
            if (GATHER) {
              b.x[s] = x[u]; b.y[s] = y[u]; b.z[s] = z[u];
              b.vx[s] = vx[u]; b.vy[s] = vy[u]; b.vz[s] = vz[u];
              b.pid[s] = pid[u];
              b.cell[s] = cell_of(s);
              if (weighted) b.w[s] = w[u];
              if (src) src[s] = q[u];
              if (SUMS) sacc += mean_rel_speed((double)vx[u], (double)vy[u], (double)vz[u], n0x, n0y, n0z, vnT);
            }
            DepRec<T> r;
            if (ABS) {
              int i;
              double f;
              cell_coord((double)x[u], g.p0[0], g.d[0], g.nx, i, f); r.fx = (T)f;
              cell_coord((double)y[u], g.p0[1], g.d[1], g.ny, i, f); r.fy = (T)f;
              cell_coord((double)z[u], g.p0[2], g.d[2], g.nz, i, f); r.fz = (T)f;
            } else {
              r.fx = x[u]; r.fy = y[u]; r.fz = z[u];
            }
            r.vx = vx[u]; r.vy = vy[u]; r.vz = vz[u];
            r.w = w[u];
            r.pad = (T)0;
            rec[e] = r;
          }
        }
      }
    };
    const int wave0 = (int)(threadIdx.x & ~63u);
    const int passes = m > wave0 ? (m - wave0 + 255) / 256 : 0;
    static_assert(kDepUnroll == 3, "one staging variant per pass count");
    if (passes >= 3) stage(std::integral_constant<int, 3>());
    else if (passes == 2) stage(std::integral_constant<int, 2>());
    else if (passes == 1) stage(std::integral_constant<int, 1>());
    __syncthreads();
    const uint32_t a0 = lo > base ? lo : base;
    const uint32_t a1 = hi < base + m ? hi : base + m;
    // the cell's ions in slot order; records read four at a time, summed in order
    auto add = [&](const DepRec<T> &r) {
      const T wx = (corner & 1) ? r.fx : (T)1 - r.fx;
      const T wy = (corner & 2) ? r.fy : (T)1 - r.fy;
      const T wz = (corner & 4) ? r.fz : (T)1 - r.fz;
      const T W = (wz * wy) * wx;
      const T a = weighted ? r.w * W : W;
      acc[0] += a;
      acc[1] = fmad(a, r.vx, acc[1]);
      acc[2] = fmad(a, r.vy, acc[2]);
      acc[3] = fmad(a, r.vz, acc[3]);
    };
    uint32_t s = a0;
    for (; s + 4 <= a1; s += 4) {
      const DepRec<T> r0 = rec[s - base], r1 = rec[s + 1 - base], r2 = rec[s + 2 - base], r3 = rec[s + 3 - base];
      add(r0); add(r1); add(r2); add(r3);
    }
    for (; s < a1; s++) add(rec[s - base]);
    __syncthreads();
  }
  if constexpr (XC) {
    // every block holds kDepCells cells (nc is a multiple of it)
    __shared__ T Ps[kDepCells][8][4];
#pragma unroll
    for (int mm = 0; mm < 4; mm++) Ps[cl][corner][mm] = acc[mm];
    __syncthreads();
    const int64_t lb = c0 / kDepCells, nbk = g.nc / kDepCells;
    T *E = P + (int64_t)16 * g.nc;
    for (int e = (int)threadIdx.x; e < 16 * kDepCells; e += 256) {
      const int rm = e / kDepCells, c = e % kDepCells, r = rm >> 2, mm = rm & 3;
      T v = Ps[c][2 * r][mm];
      if (c > 0) v = v + Ps[c - 1][2 * r + 1][mm];
      P[(int64_t)rm * g.nc + c0 + c] = v;
      if (c == kDepCells - 1) E[(int64_t)rm * nbk + lb] = Ps[c][2 * r + 1][mm];
    }
  } else if (own) {
#pragma unroll
    for (int mm = 0; mm < 4; mm++) P[(int64_t)(4 * corner + mm) * g.nc + c0 + cl] = acc[mm];
  }
  if (GATHER && SUMS) {
    const double t = block_sum(sacc, sh);
    if (threadIdx.x == 0) sred[blockIdx.x] = t;
  }
}

// Node moments: sum over the 8 cells sharing the node, corner order q = 0..7.
template <typename T>
__global__ __launch_bounds__(256) void k_node_sum(Grid g, const T *__restrict__ P, T *__restrict__ n, T *__restrict__ gx,
                                                  T *__restrict__ gy, T *__restrict__ gz)
{
  const int64_t c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  int i, j, k;
  split_cell(g, (int)c, i, j, k);
  T m[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int64_t s = lin(g, i - (q & 1), j - ((q >> 1) & 1), k - (q >> 2));
#pragma unroll
    for (int a = 0; a < 4; a++) m[a] += P[(int64_t)(4 * q + a) * g.nc + s];
  }
  n[c] = m[0]; gx[c] = m[1]; gy[c] = m[2]; gz[c] = m[3];
}

// Node moments from the x-combined partials of k_cell_partials<XC>: the four
// y-z corners r = 0..3 in order, each the x pair made in the block, plus for
// nodes at a block's first x position the x-high corners of the cell before
// it (the previous block's last cell, or the row's last across the periodic
// wrap), added as the block adds them.
template <typename T>
__global__ __launch_bounds__(256) void k_node_sum_xc(Grid g, const T *__restrict__ P, T *__restrict__ n,
                                                     T *__restrict__ gx, T *__restrict__ gy, T *__restrict__ gz)
{
  const int64_t c = xcd_block() * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  int i, j, k;
  split_cell(g, (int)c, i, j, k);
  const int64_t nbk = g.nc / kDepCells;
  const T *E = P + (int64_t)16 * g.nc;
  const bool edge = i % kDepCells == 0;
  // every lane loads an E value (block 0's for the others, unused), so the
  // edge lanes' loads go out with the rest instead of after them
  T pv[16], ev[16];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int jj = j - (r & 1), kk = k - (r >> 1);
    const int64_t s = lin(g, i, jj, kk);
    const int64_t bw = edge ? lin(g, i - 1, jj, kk) / kDepCells : 0;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      pv[4 * r + a] = P[(int64_t)(4 * r + a) * g.nc + s];
      ev[4 * r + a] = E[(int64_t)(4 * r + a) * nbk + bw];
    }
  }
  T m[4] = {(T)0, (T)0, (T)0, (T)0};
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const T v = edge ? pv[4 * r + a] + ev[4 * r + a] : pv[4 * r + a];
      m[a] += v;
    }
  }
  n[c] = m[0]; gx[c] = m[1]; gy[c] = m[2]; gz[c] = m[3];
}

// ---- device initialisation (seeded; mirrors the oracle's placement) ----------------
template <bool MIXED>
__global__ __launch_bounds__(256) void k_init(Grid g, int64_t ppc, uint32_t k0, uint32_t k1, double v0x, double v0y,
                                              double v0z, double vTx, double vTy, double vTz, int pert, double amp,
                                              PState<double> sd, PState<float> sf)
{
  const int64_t np = g.nc * ppc;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int64_t c = p / ppc;
  const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((int64_t)g.nx * g.ny));
  const U4 a = philox4x32_10(U4{(uint32_t)p, 0u, STREAM_INIT, 0u}, k0, k1);
  const U4 b = philox4x32_10(U4{(uint32_t)p, 0u, STREAM_INIT, 1u}, k0, k1);
  const double u0 = u24<double>(a.x), u1 = u24<double>(a.y), u2 = u24<double>(a.z);
  const double r1 = sqrt(-2.0 * log(u24<double>(b.x))), a1 = 2.0 * M_PI * u24<double>(b.y);
  const double r2 = sqrt(-2.0 * log(u24<double>(b.z))), a2 = 2.0 * M_PI * u24<double>(b.w);
  const double vx = v0x + vTx * (r1 * cos(a1));
  const double vy = v0y + vTy * (r1 * sin(a1));
  const double vz = v0z + vTz * (r2 * cos(a2));
  double wgt = 1.0;
  if (pert == CHY_PERT_UNIFORM || pert == CHY_PERT_GAUSSIAN) {
    const U4 wv = philox4x32_10(U4{(uint32_t)c, 0u, STREAM_WEIGHT, 0u}, k0, k1);
    const double w0 = u24<double>(wv.x), w1 = u24<double>(wv.y);
    wgt = (pert == CHY_PERT_UNIFORM) ? 1.0 + amp * (2.0 * w0 - 1.0)
                                     : 1.0 + amp * (sqrt(-2.0 * log(w0)) * cos(2.0 * M_PI * w1));
  }
  if (!MIXED) {
    // L = d * N exactly as the oracle's init computes it
    const double Lx = g.d[0] * (double)g.nx, Ly = g.d[1] * (double)g.ny, Lz = g.d[2] * (double)g.nz;
    const double x = wrap_pos(g.p0[0] + ((double)i + u0) * g.d[0], g.p0[0], g.p0[0] + Lx, Lx);
    const double y = wrap_pos(g.p0[1] + ((double)j + u1) * g.d[1], g.p0[1], g.p0[1] + Ly, Ly);
    const double z = wrap_pos(g.p0[2] + ((double)k + u2) * g.d[2], g.p0[2], g.p0[2] + Lz, Lz);
    sd.x[p] = x; sd.y[p] = y; sd.z[p] = z;
    sd.vx[p] = vx; sd.vy[p] = vy; sd.vz[p] = vz;
    sd.pid[p] = (uint32_t)p;
    if (pert == CHY_PERT_SINUSOIDAL_X) wgt = 1.0 + amp * sin(2.0 * M_PI * (x - g.p0[0]) / Lx);
    if (sd.w) sd.w[p] = wgt;
  } else {
    if (pert == CHY_PERT_SINUSOIDAL_X) wgt = 1.0 + amp * sin(2.0 * M_PI * ((double)i + u0) / (double)g.nx);
    // offsets exact (24-bit uniforms); the weight is 1 without a perturbation
    sf.pos[p] = Vec4<float>{(float)u0, (float)u1, (float)u2, (float)wgt};
    sf.vel[p] = Vec4<float>{(float)vx, (float)vy, (float)vz, __uint_as_float((uint32_t)p)};
    sf.cell[p] = (int32_t)c;
  }
}

// ---- the reference's initial densities (particles.c:26-45, distributions.c) --------
// Direction integers of the first three dimensions of the reference's Sobol
// generator (random.c:234-290), scaled to 30 bits: V[j][k] for bit j, dimension k.
struct SobolDirs { uint32_t V[30][3]; };

// 53-bit uniform in [0, 1) from two Philox words.
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

constexpr uint32_t STREAM_PLACE = 4;

// Ion p of a reference placement: flat-sobol (point p of the Sobol sequence,
// the call order of random.c:273-287 given in closed form: the running XOR
// after call p is the XOR of the direction integers of the set bits of the
// Gray code of p + 1; positions as distributions.c:124-136), uniform (cell
// centres of all but the last cell along each axis, cell index k fastest,
// distributions.c:71-86) or sinusoidal (rejection against 1 + 0.25 sin(2 pi x)
// with the density's maximum 1.25, distributions.c:144-232; the candidates of
// ion p from its own Philox blocks, the last one kept after 64 rejections).
// Velocities and weights as k_init.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_init_density(Grid g, int kind, int64_t np, SobolDirs sob, uint32_t k0,
                                                      uint32_t k1, double v0x, double v0y, double v0z, double vTx,
                                                      double vTy, double vTz, int pert, double amp,
                                                      PState<double> sd, PState<float> sf)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double q[3];
  if (kind == CHY_DENSITY_FLAT_SOBOL) {
    const uint32_t gc = (uint32_t)(p + 1) ^ (uint32_t)((p + 1) >> 1);
    uint32_t ix[3] = {0u, 0u, 0u};
#pragma unroll 1
    for (int j = 0; j < 30; j++) {
      if ((gc >> j) & 1u) {
        ix[0] ^= sob.V[j][0]; ix[1] ^= sob.V[j][1]; ix[2] ^= sob.V[j][2];
      }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double r = (double)ix[a] * (1.0 / 1073741824.0);
      double v = g.p0[a] + (r * (g.L[a] + g.d[a]) - 0.5 * g.d[a]);
      if (v < g.p0[a]) v = g.p0[a] + 0.01 * g.d[a];
      else if (v > g.p1[a]) v = g.p1[a] - 0.01 * g.d[a];
      q[a] = wrap_pos(v, g.p0[a], g.p1[a], g.L[a]);
    }
  } else if (kind == CHY_DENSITY_UNIFORM) {
    const int64_t mx = g.nx - 1, my = g.ny - 1, mz = g.nz - 1, m = mx * my * mz;
    const int64_t r = p % m;
    const int64_t k = r % mz, j = (r / mz) % my, i = r / (mz * my);
    q[0] = g.p0[0] + g.d[0] * ((double)(i + 1) - 0.5);
    q[1] = g.p0[1] + g.d[1] * ((double)(j + 1) - 0.5);
    q[2] = g.p0[2] + g.d[2] * ((double)(k + 1) - 0.5);
  } else {
    q[0] = g.p0[0]; q[1] = g.p0[1]; q[2] = g.p0[2];
#pragma unroll 1
    for (uint32_t a = 0; a < 64; a++) {
      const U4 b0 = philox4x32_10(U4{(uint32_t)p, a, STREAM_PLACE, 0u}, k0, k1);
      const U4 b1 = philox4x32_10(U4{(uint32_t)p, a, STREAM_PLACE, 1u}, k0, k1);
      const double u[3] = {u53(b0.x, b0.y), u53(b0.z, b0.w), u53(b1.x, b1.y)};
#pragma unroll
      for (int e = 0; e < 3; e++) q[e] = wrap_pos(g.p0[e] + u[e] * g.L[e], g.p0[e], g.p1[e], g.L[e]);
      const double dens = 1.0 + 0.25 * sin(2.0 * M_PI * q[0]);
      if (dens > u24<double>(b1.z) * 1.25) break;
    }
  }
  const U4 b = philox4x32_10(U4{(uint32_t)p, 0u, STREAM_INIT, 1u}, k0, k1);
  const double r1 = sqrt(-2.0 * log(u24<double>(b.x))), a1 = 2.0 * M_PI * u24<double>(b.y);
  const double r2 = sqrt(-2.0 * log(u24<double>(b.z))), a2 = 2.0 * M_PI * u24<double>(b.w);
  const double vx = v0x + vTx * (r1 * cos(a1));
  const double vy = v0y + vTy * (r1 * sin(a1));
  const double vz = v0z + vTz * (r2 * cos(a2));
  // cell of the position (the clamped truncation of cell_coord)
  int ci[3];
  double fr[3];
  cell_coord(q[0], g.p0[0], g.d[0], g.nx, ci[0], fr[0]);
  cell_coord(q[1], g.p0[1], g.d[1], g.ny, ci[1], fr[1]);
  cell_coord(q[2], g.p0[2], g.d[2], g.nz, ci[2], fr[2]);
  const int64_t c = ci[0] + (int64_t)g.nx * (ci[1] + (int64_t)g.ny * ci[2]);
  double wgt = 1.0;
  if (pert == CHY_PERT_SINUSOIDAL_X) {
    wgt = 1.0 + amp * sin(2.0 * M_PI * (q[0] - g.p0[0]) / (g.d[0] * (double)g.nx));
  } else if (pert == CHY_PERT_UNIFORM || pert == CHY_PERT_GAUSSIAN) {
    const U4 wv = philox4x32_10(U4{(uint32_t)c, 0u, STREAM_WEIGHT, 0u}, k0, k1);
    const double w0 = u24<double>(wv.x), w1 = u24<double>(wv.y);
    wgt = (pert == CHY_PERT_UNIFORM) ? 1.0 + amp * (2.0 * w0 - 1.0)
                                     : 1.0 + amp * (sqrt(-2.0 * log(w0)) * cos(2.0 * M_PI * w1));
  }
  if (!MIXED) {
    sd.x[p] = q[0]; sd.y[p] = q[1]; sd.z[p] = q[2];
    sd.vx[p] = vx; sd.vy[p] = vy; sd.vz[p] = vz;
    sd.pid[p] = (uint32_t)p;
    if (sd.w) sd.w[p] = wgt;
  } else {
    // cell + FP32 offset as an upload of the FP64 position makes them
    const int nn[3] = {g.nx, g.ny, g.nz};
    float o[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float f = (float)fr[a];
      if (f >= 1.0f) { f = 0.0f; ci[a] = (ci[a] + 1) % nn[a]; }
      o[a] = f;
    }
    sf.pos[p] = Vec4<float>{o[0], o[1], o[2], (float)wgt};
    sf.vel[p] = Vec4<float>{(float)vx, (float)vy, (float)vz, __uint_as_float((uint32_t)p)};
    sf.cell[p] = (int32_t)(ci[0] + g.nx * (ci[1] + g.ny * ci[2]));
  }
}

// Direction integers of the reference's Sobol generator (random.c:241-270):
// degrees 1, 2, 3, polynomial codes 0, 1, 1, leading odd integers {1},
// {1, 1}, {1, 3, 7}, and the recurrence as random.c:260-269 writes it (one
// XOR of the previous integer per set polynomial bit).
static SobolDirs sobol_dirs()
{
  SobolDirs s;
  const int deg[3] = {1, 2, 3};
  const unsigned code[3] = {0u, 1u, 1u};
  const uint32_t lead[3][3] = {{1u, 0u, 0u}, {1u, 1u, 0u}, {1u, 3u, 7u}};
  for (int k = 0; k < 3; k++) {
    const int dg = deg[k];
    for (int j = 0; j < dg; j++) s.V[j][k] = lead[k][j] << (29 - j);
    for (int j = dg; j < 30; j++) {
      uint32_t v = s.V[j - dg][k];
      v ^= v >> dg;
      unsigned bits = code[k];
      for (int l = dg - 1; l >= 1; l--) {
        if (bits & 1u) v ^= s.V[j - 1][k];
        bits >>= 1;
      }
      s.V[j][k] = v;
    }
  }
  return s;
}

// ---- host orchestration ------------------------------------------------------------
static unsigned grid_for(int64_t n, unsigned cap = 4096)
{
  const int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

// A placement kernel (k_perm, k_scatter) computed a slot outside the ion
// range on an earlier sort: the ion order is not a permutation, so the state
// is lost.  The flag is read with the next sort's other read-backs.
// Status words to the host: out[q] = *w[q] for the 32-bit words given
// (q < 4), then out[4 + q] = u[q] for q < nu; out is coherent host memory.
// Last, the launch's number s_no goes to the sequence word out[kSeqWord].
__global__ __launch_bounds__(64) void k_status(const uint32_t *w0, const uint32_t *w1, const uint32_t *w2,
                                               const uint32_t *w3, const unsigned long long *u, int nu,
                                               unsigned long long *out, unsigned long long s_no)
{
  const int t = (int)threadIdx.x;
  const uint32_t *w = t == 0 ? w0 : t == 1 ? w1 : t == 2 ? w2 : w3;
  if (t < 4 && w) out[t] = *w;
  if (t < nu) out[4 + t] = u[t];
  // every writer's values reach the host before the number does
  __threadfence_system();
  __syncthreads();
  if (t == 0) *(volatile unsigned long long *)(out + kSeqWord) = s_no;
}

static int push_finish(chy_ctx *ctx, const unsigned long long cs[6]);

// Reads the words (null entries read 0) and nu counters in one
// synchronisation: w[q] into v32[q], the counters into v64.  A push whose
// counters are still unread (ctx->push_pending) has them read in the same
// synchronisation and checked first.
static int read_status(chy_ctx *ctx, const uint32_t *const w[4], uint32_t v32[4], const unsigned long long *u, int nu,
                       unsigned long long *v64)
{
  const bool push = ctx->push_pending;
  if (nu > kSeqWord - 4 || (push && nu > 0)) return set_err(ctx, CHY_ERR_LIB, "read_status: bad counter request");
  if (push) { u = ctx->cstats; nu = 6; }
  const unsigned long long s_no = next_seq(ctx);
  hipLaunchKernelGGL(k_status, 1, 64, 0, ctx->stream, w[0], w[1], w[2], w[3], u, nu,
                     (unsigned long long *)ctx->status_dev, s_no);
  TRY(wait_seq(ctx, s_no));
  for (int q = 0; q < 4; q++) v32[q] = w[q] ? (uint32_t)ctx->h_status[q] : 0u;
  if (push) {
    unsigned long long cs[6];
    for (int q = 0; q < 6; q++) cs[q] = ctx->h_status[4 + q];
    return push_finish(ctx, cs);
  }
  for (int q = 0; q < nu; q++) v64[q] = ctx->h_status[4 + q];
  return CHY_OK;
}

// The mixed path deposits through the x-combined partials when whole blocks
// of cells lie in one x row.
static bool x_combined(const chy_ctx *ctx) { return ctx->mixed && ctx->g.nx % kDepCells == 0; }

static int lost_placement(chy_ctx *ctx)
{
  launch_fill<uint32_t>(ctx, 1, ctx->sortflag + 1, 0u);
  return set_err(ctx, CHY_ERR_LIB, "sort: an earlier placement put an ion outside the ion range; the particle state is "
                                   "not a permutation of the ions");
}

// The neighbour-move sort (see k_trans); false when it does not apply and the
// general sort must run (counts[] is then filled here).
static int sort_neighbour(chy_ctx *ctx, bool &done)
{
  done = false;
  const int64_t nc = ctx->g.nc, np = ctx->np;
  const int64_t nt = (nc + kScanTile - 1) / kScanTile;
  const int32_t *cell = ctx->mixed ? ctx->pf[ctx->cur].cell : ctx->pd[ctx->cur].cell;
  const unsigned ncb = (unsigned)((nc + 255) / 256);
  launch_fill<uint32_t>(ctx, 1, ctx->sortflag, 0u);  // word 1 (placements lost) stays until read
  launch_fill<uint32_t>(ctx, 1, ctx->maxcount, 0u);
  // the group sizes and ranks from the push's move directions (k_trans_dir,
  // its LDS staging sized from the mean occupancy) or from the new keys (k_trans)
  const bool fromdir = ctx->dir_valid;
  const bool dense = np > 20 * nc;
  timer_start(ctx, "sort_trans");
  if (fromdir) {
    if (dense)
      hipLaunchKernelGGL((k_trans_dir<16>), ncb, 256, 0, ctx->stream, ctx->g, ctx->offsets,
                         (const uint8_t *)ctx->dir, ctx->info, ctx->trans, ctx->sortflag);
    else
      hipLaunchKernelGGL((k_trans_dir<8>), ncb, 256, 0, ctx->stream, ctx->g, ctx->offsets,
                         (const uint8_t *)ctx->dir, ctx->info, ctx->trans, ctx->sortflag);
  } else {
    hipLaunchKernelGGL(k_trans, ncb, 256, 0, ctx->stream, ctx->g, ctx->offsets, cell, ctx->info, ctx->trans,
                       ctx->sortflag);
  }
  hipLaunchKernelGGL(k_counts, ncb, 256, 0, ctx->stream, ctx->g, ctx->trans, ctx->counts, ctx->nbase);
  timer_stop(ctx, "sort_trans");
  timer_start(ctx, "sort_scan");
  hipLaunchKernelGGL(k_tile_sum, (unsigned)nt, kBlock, 0, ctx->stream, nc, ctx->counts, ctx->scan_tmp,
                     ctx->scan_tmp + nt + 1);
  hipLaunchKernelGGL(k_scan_tiles, 1, kBlock, 0, ctx->stream, nt, ctx->scan_tmp, ctx->scan_tmp + nt + 1, ctx->maxcount);
  hipLaunchKernelGGL(k_tile_scan, (unsigned)nt, kBlock, 0, ctx->stream, nc, ctx->counts, ctx->scan_tmp, ctx->offs_alt,
                     ctx->cursor, (uint32_t)np);
  timer_stop(ctx, "sort_scan");
  uint32_t v[4];
  {
    const uint32_t *w[4] = {ctx->sortflag, ctx->sortflag + 1, ctx->maxcount, ctx->scan_tmp + nt};
    TRY(read_status(ctx, w, v, nullptr, 0, nullptr));
  }
  const uint32_t flag = v[0], lost = v[1], mx = v[2], total = v[3];
  if (lost) return lost_placement(ctx);
  if (flag || mx > kMaxCellOccupancy || total != (uint32_t)np) {
    launch_fill<uint32_t>(ctx, nc + 1, ctx->counts, 0u);
    hipLaunchKernelGGL(k_keys_count, grid_for(np), kBlock, 0, ctx->stream, np, cell, ctx->counts);
    return CHY_OK;
  }
  const unsigned nb = grid_for(np);
  timer_start(ctx, "sort_scatter");
  hipLaunchKernelGGL(k_perm, nb, kBlock, 0, ctx->stream, ctx->g, np, cell, ctx->info, ctx->nbase, ctx->offs_alt,
                     ctx->perm, ctx->sortflag + 1);
  timer_stop(ctx, "sort_scatter");
  const chy_params &p = ctx->prm;
  const bool sums = p.mn > 0.0;  // the collision normaliser is needed
  const unsigned nbc = (unsigned)((nc + kDepCells - 1) / kDepCells);
  timer_start(ctx, "sort_reorder");
#define GATHER(T, ABS, S, XC, st)                                                                                 \
  hipLaunchKernelGGL((k_cell_partials<T, ABS, true, S, XC>), nbc, kBlock, 0, ctx->stream, ctx->g, ctx->offs_alt,   \
                     st[ctx->cur], ctx->weighted, (T *)ctx->partials, ctx->perm, st[1 - ctx->cur], p.vn0[0],      \
                     p.vn0[1], p.vn0[2], p.vnT, ctx->sred, ctx->sort_src)
  if (ctx->mixed) {
    if (x_combined(ctx)) { if (sums) GATHER(float, false, true, true, ctx->pf); else GATHER(float, false, false, true, ctx->pf); }
    else { if (sums) GATHER(float, false, true, false, ctx->pf); else GATHER(float, false, false, false, ctx->pf); }
  } else { if (sums) GATHER(double, true, true, false, ctx->pd); else GATHER(double, true, false, false, ctx->pd); }
#undef GATHER
  timer_stop(ctx, "sort_reorder");
  if (sums) launch_combine_sum(ctx, ctx->sred, (int)nbc, 1, ctx->S_dev);
  ctx->partials_fresh = true;  // the deposit's cell partials of the new order are made
  std::swap(ctx->offsets, ctx->offs_alt);
  ctx->cur = 1 - ctx->cur;
  ctx->dir_valid = false;  // the order is new: directions and keys agree with it
  ctx->keys_moved = false;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, std::string("sort: ") + hipGetErrorString(e));
  done = true;
  return CHY_OK;
}

int particles_sort(chy_ctx *ctx)
{
  if (ctx->order_valid && ctx->np > 0) {
    bool done = false;
    TRY(sort_neighbour(ctx, done));
    if (done) return CHY_OK;
  }
  ctx->order_valid = false;
  ctx->dir_valid = false;
  ctx->keys_moved = false;
  ctx->partials_fresh = false;
  // counts[] was filled by the key pass (push, keys kernel or the fallback above)
  const int64_t nc = ctx->g.nc, np = ctx->np;
  const int64_t nt = (nc + kScanTile - 1) / kScanTile;
  launch_fill<uint32_t>(ctx, 1, ctx->maxcount, 0u);
  timer_start(ctx, "sort_scan");
  hipLaunchKernelGGL(k_tile_sum, (unsigned)nt, kBlock, 0, ctx->stream, nc, ctx->counts, ctx->scan_tmp,
                     ctx->scan_tmp + nt + 1);
  hipLaunchKernelGGL(k_scan_tiles, 1, kBlock, 0, ctx->stream, nt, ctx->scan_tmp, ctx->scan_tmp + nt + 1, ctx->maxcount);
  hipLaunchKernelGGL(k_tile_scan, (unsigned)nt, kBlock, 0, ctx->stream, nc, ctx->counts, ctx->scan_tmp, ctx->offsets,
                     ctx->cursor, (uint32_t)np);
  timer_stop(ctx, "sort_scan");
  // the stable per-cell rank is quadratic in the cell occupancy: bound it; and
  // the histogram must account for every ion before the scatter uses it
  uint32_t v[4];
  {
    const uint32_t *w[4] = {ctx->maxcount, ctx->scan_tmp + nt, ctx->sortflag + 1, nullptr};
    TRY(read_status(ctx, w, v, nullptr, 0, nullptr));
  }
  const uint32_t mx = v[0], total = v[1], lost = v[2];
  if (lost) return lost_placement(ctx);
  if (total != (uint32_t)np)
    return set_err(ctx, CHY_ERR_LIB, "sort: the cell histogram holds " + std::to_string(total) + " ions, not " +
                                         std::to_string(np));
  if (mx > kMaxCellOccupancy)
    return set_err(ctx, CHY_ERR_SUP, "a cell holds " + std::to_string(mx) + " particles; at most " +
                                         std::to_string(kMaxCellOccupancy) + " are supported");
  const int32_t *cell = ctx->mixed ? ctx->pf[ctx->cur].cell : ctx->pd[ctx->cur].cell;
  timer_start(ctx, "sort_scatter");
  hipLaunchKernelGGL(k_scatter, grid_for(np), kBlock, 0, ctx->stream, np, cell, ctx->cursor, ctx->perm, ctx->slotkey,
                     ctx->sortflag + 1);
  timer_stop(ctx, "sort_scatter");
  const unsigned nb = grid_for(np);
  const chy_params &p = ctx->prm;
  const bool sums = p.mn > 0.0;  // the collision normaliser is needed
  timer_start(ctx, "sort_reorder");
#define REORDER(T, S, st)                                                                                  \
  hipLaunchKernelGGL((k_reorder<T, S>), nb, kBlock, 0, ctx->stream, np, ctx->perm, ctx->slotkey, ctx->offsets, \
                     st[ctx->cur], st[1 - ctx->cur], ctx->weighted, p.vn0[0], p.vn0[1], p.vn0[2], p.vnT, ctx->sred, \
                     ctx->sort_src)
  if (ctx->mixed) { if (sums) REORDER(float, true, ctx->pf); else REORDER(float, false, ctx->pf); }
  else { if (sums) REORDER(double, true, ctx->pd); else REORDER(double, false, ctx->pd); }
#undef REORDER
  timer_stop(ctx, "sort_reorder");
  if (sums) launch_combine_sum(ctx, ctx->sred, (int)nb, 1, ctx->S_dev);
  ctx->cur = 1 - ctx->cur;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, std::string("sort: ") + hipGetErrorString(e));
  ctx->order_valid = true;
  ctx->dir_valid = false;
  ctx->keys_moved = false;
  return CHY_OK;
}

int particles_deposit(chy_ctx *ctx)
{
  const unsigned nb = (unsigned)((ctx->g.nc + kBlock - 1) / kBlock);
  const unsigned nbc = (unsigned)((ctx->g.nc + kDepCells - 1) / kDepCells);
  // after the neighbour-move sort the cell partials are already made
  const bool cells = !ctx->partials_fresh;
  ctx->partials_fresh = false;
  if (ctx->mixed) {
    float *P = (float *)ctx->partials;
    const bool xc = x_combined(ctx);
    if (cells) {
      timer_start(ctx, "deposit_cells");
      if (xc)
        hipLaunchKernelGGL((k_cell_partials<float, false, false, false, true>), nbc, kBlock, 0, ctx->stream, ctx->g,
                           ctx->offsets, ctx->pf[ctx->cur], ctx->weighted, P, (const uint32_t *)nullptr,
                           PState<float>(), 0.0, 0.0, 0.0, 0.0, (double *)nullptr, (uint32_t *)nullptr);
      else
        hipLaunchKernelGGL((k_cell_partials<float, false, false, false>), nbc, kBlock, 0, ctx->stream, ctx->g,
                           ctx->offsets, ctx->pf[ctx->cur], ctx->weighted, P, (const uint32_t *)nullptr,
                           PState<float>(), 0.0, 0.0, 0.0, 0.0, (double *)nullptr, (uint32_t *)nullptr);
      timer_stop(ctx, "deposit_cells");
    }
    timer_start(ctx, "deposit_nodes");
    if (xc)
      hipLaunchKernelGGL((k_node_sum_xc<float>), nb, kBlock, 0, ctx->stream, ctx->g, P, (float *)ctx->mom[0],
                         (float *)ctx->mom[1], (float *)ctx->mom[2], (float *)ctx->mom[3]);
    else
      hipLaunchKernelGGL((k_node_sum<float>), nb, kBlock, 0, ctx->stream, ctx->g, P, (float *)ctx->mom[0],
                         (float *)ctx->mom[1], (float *)ctx->mom[2], (float *)ctx->mom[3]);
    timer_stop(ctx, "deposit_nodes");
  } else {
    double *P = (double *)ctx->partials;
    if (cells) {
      timer_start(ctx, "deposit_cells");
      hipLaunchKernelGGL((k_cell_partials<double, true, false, false>), nbc, kBlock, 0, ctx->stream, ctx->g, ctx->offsets,
                         ctx->pd[ctx->cur], ctx->weighted, P, (const uint32_t *)nullptr, PState<double>(), 0.0, 0.0, 0.0, 0.0, (double *)nullptr, (uint32_t *)nullptr);
      timer_stop(ctx, "deposit_cells");
    }
    timer_start(ctx, "deposit_nodes");
    hipLaunchKernelGGL((k_node_sum<double>), nb, kBlock, 0, ctx->stream, ctx->g, P, (double *)ctx->mom[0],
                       (double *)ctx->mom[1], (double *)ctx->mom[2], (double *)ctx->mom[3]);
    timer_stop(ctx, "deposit_nodes");
  }
  ctx->moments_valid = true;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(ctx, CHY_ERR_LIB, std::string("deposit: ") + hipGetErrorString(e));
  return CHY_OK;
}

void deposit_state(chy_ctx *ctx, const PState<double> &st, const uint32_t *offs, bool weighted, double *const out[4])
{
  const unsigned nb = (unsigned)((ctx->g.nc + kBlock - 1) / kBlock);
  const unsigned nbc = (unsigned)((ctx->g.nc + kDepCells - 1) / kDepCells);
  double *P = (double *)ctx->partials;
  hipLaunchKernelGGL((k_cell_partials<double, true, false, false>), nbc, kBlock, 0, ctx->stream, ctx->g, offs, st, weighted, P, (const uint32_t *)nullptr, PState<double>(), 0.0, 0.0, 0.0, 0.0, (double *)nullptr, (uint32_t *)nullptr);
  hipLaunchKernelGGL((k_node_sum<double>), nb, kBlock, 0, ctx->stream, ctx->g, P, out[0], out[1], out[2], out[3]);
}

int particles_normaliser(chy_ctx *ctx)
{
  const chy_params &p = ctx->prm;
  const unsigned nb = grid_for(ctx->np);
  if (ctx->mixed) {
    hipLaunchKernelGGL(k_meanrel_mixed, nb, kBlock, 0, ctx->stream, ctx->np, ctx->pf[ctx->cur].vel, p.vn0[0], p.vn0[1],
                       p.vn0[2], p.vnT, ctx->sred);
  } else {
    const PState<double> &s = ctx->pd[ctx->cur];
    hipLaunchKernelGGL((k_meanrel<double>), nb, kBlock, 0, ctx->stream, ctx->np, s.vx, s.vy, s.vz, p.vn0[0],
                       p.vn0[1], p.vn0[2], p.vnT, ctx->sred);
  }
  launch_combine_sum(ctx, ctx->sred, (int)nb, 1, ctx->S_dev);
  return CHY_OK;
}

// Keys + histogram + sort + normaliser after new particle data.
static int settle(chy_ctx *ctx)
{
  ctx->order_valid = false;  // new data: the general sort
  ctx->dir_valid = false;
  ctx->keys_moved = false;
  ctx->partials_fresh = false;
  launch_fill<uint32_t>(ctx, ctx->g.nc + 1, ctx->counts, 0u);
  if (ctx->mixed)
    hipLaunchKernelGGL(k_keys_count, grid_for(ctx->np), kBlock, 0, ctx->stream, ctx->np, ctx->pf[ctx->cur].cell,
                       ctx->counts);
  else
    hipLaunchKernelGGL(k_keys_fp64, grid_for(ctx->np), kBlock, 0, ctx->stream, ctx->g, ctx->np, ctx->pd[ctx->cur],
                       ctx->counts);
  int rc = particles_sort(ctx);
  if (rc) return rc;
  return particles_normaliser(ctx);
}

// The device wrap rule on the host (upload of positions from the caller).
static double host_wrap(double x, double p0, double p1, double L)
{
  if (!(x >= p0 - L && x < p1 + L)) {
    if (!std::isfinite(x)) return x;
    x = x - std::floor((x - p0) / L) * L;
  }
  if (x < p0) { x += L; if (x >= p1) x = p0; }
  else if (x >= p1) { x -= L; if (x < p0) x = p0; }
  return x;
}

int particles_upload(chy_ctx *ctx, int64_t np, const double *pos, const double *vel, const double *w,
                     const uint32_t *pid)
{
  if (np < 0 || np > ctx->cap) return set_err(ctx, CHY_ERR_ARG_RANGE, "particle count exceeds the context capacity");
  if (w && !ctx->weighted) return set_err(ctx, CHY_ERR_ARG_INCOMP, "weights given but the context has no weight buffers");
  ctx->np = np;
  if (np == 0) return settle(ctx);
  const chy::Grid &g = ctx->g;
  std::vector<uint32_t> hp(np);
  for (int64_t i = 0; i < np; i++) hp[i] = pid ? pid[i] : (uint32_t)i;
  if (!ctx->mixed) {
    std::vector<double> a(np);
    PState<double> &s = ctx->pd[ctx->cur];
    double *dst[6] = {s.x, s.y, s.z, s.vx, s.vy, s.vz};
    for (int f = 0; f < 6; f++) {
      const double *src = f < 3 ? pos : vel;
      for (int64_t i = 0; i < np; i++) a[i] = src[4 * i + (f % 3)];
      if (f < 3) {
        // wrap into [p0, p1) as the move does
        for (int64_t i = 0; i < np; i++) {
          a[i] = host_wrap(a[i], g.p0[f], g.p1[f], g.L[f]);
        }
      }
      CHY_HIP(hipMemcpyAsync(dst[f], a.data(), sizeof(double) * np, hipMemcpyHostToDevice, ctx->stream));
      CHY_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (w) CHY_HIP(hipMemcpyAsync(s.w, w, sizeof(double) * np, hipMemcpyHostToDevice, ctx->stream));
    else if (ctx->weighted) launch_fill<double>(ctx, np, s.w, 1.0);
    CHY_HIP(hipMemcpyAsync(s.pid, hp.data(), sizeof(uint32_t) * np, hipMemcpyHostToDevice, ctx->stream));
    CHY_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    PState<float> &s = ctx->pf[ctx->cur];
    std::vector<Vec4<float>> hp4(np), hv4(np);
    std::vector<int32_t> cell(np);
    const int nn[3] = {g.nx, g.ny, g.nz};
    for (int64_t i = 0; i < np; i++) {
      int ci[3];
      float o[3];
      for (int a = 0; a < 3; a++) {
        const double p0 = g.p0[a];
        const double x = host_wrap(pos[4 * i + a], p0, g.p1[a], g.L[a]);
        const double xi = (x - p0) / g.d[a];
        int ii = (int)xi;
        if (ii > nn[a] - 1) ii = nn[a] - 1;
        float f = (float)(xi - (double)ii);
        if (f >= 1.0f) { f = 0.0f; ii = (ii + 1) % nn[a]; }
        ci[a] = ii;
        o[a] = f;
      }
      uint32_t id = hp[i];
      float idf;
      memcpy(&idf, &id, sizeof idf);
      hp4[i] = Vec4<float>{o[0], o[1], o[2], w ? (float)w[i] : 1.0f};
      hv4[i] = Vec4<float>{(float)vel[4 * i], (float)vel[4 * i + 1], (float)vel[4 * i + 2], idf};
      cell[i] = ci[0] + g.nx * (ci[1] + g.ny * ci[2]);
    }
    CHY_HIP(hipMemcpyAsync(s.pos, hp4.data(), sizeof(Vec4<float>) * np, hipMemcpyHostToDevice, ctx->stream));
    CHY_HIP(hipMemcpyAsync(s.vel, hv4.data(), sizeof(Vec4<float>) * np, hipMemcpyHostToDevice, ctx->stream));
    CHY_HIP(hipMemcpyAsync(s.cell, cell.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, ctx->stream));
    CHY_HIP(hipStreamSynchronize(ctx->stream));
  }
  return settle(ctx);
}

int particles_download(chy_ctx *ctx, double *pos, double *vel, double *w, uint32_t *pid)
{
  const int64_t np = ctx->np;
  if (np == 0) return CHY_OK;
  const chy::Grid &g = ctx->g;
  if (!ctx->mixed) {
    const PState<double> &s = ctx->pd[ctx->cur];
    std::vector<double> a[6];
    const double *src[6] = {s.x, s.y, s.z, s.vx, s.vy, s.vz};
    for (int f = 0; f < 6; f++) {
      a[f].resize(np);
      CHY_HIP(hipMemcpyAsync(a[f].data(), src[f], sizeof(double) * np, hipMemcpyDeviceToHost, ctx->stream));
    }
    CHY_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < np; i++) {
      if (pos) {
        for (int q = 0; q < 3; q++) pos[4 * i + q] = a[q][i];
        pos[4 * i + 3] = sqrt(a[0][i] * a[0][i] + a[1][i] * a[1][i] + a[2][i] * a[2][i]);
      }
      if (vel) {
        for (int q = 0; q < 3; q++) vel[4 * i + q] = a[3 + q][i];
        vel[4 * i + 3] = sqrt(a[3][i] * a[3][i] + a[4][i] * a[4][i] + a[5][i] * a[5][i]);
      }
    }
    if (w) {
      if (ctx->weighted) CHY_HIP(hipMemcpy(w, s.w, sizeof(double) * np, hipMemcpyDeviceToHost));
      else for (int64_t i = 0; i < np; i++) w[i] = 1.0;
    }
    if (pid) CHY_HIP(hipMemcpy(pid, s.pid, sizeof(uint32_t) * np, hipMemcpyDeviceToHost));
  } else {
    const PState<float> &s = ctx->pf[ctx->cur];
    std::vector<Vec4<float>> hp4(np), hv4(np);
    std::vector<int32_t> cell(np);
    CHY_HIP(hipMemcpyAsync(hp4.data(), s.pos, sizeof(Vec4<float>) * np, hipMemcpyDeviceToHost, ctx->stream));
    CHY_HIP(hipMemcpyAsync(hv4.data(), s.vel, sizeof(Vec4<float>) * np, hipMemcpyDeviceToHost, ctx->stream));
    CHY_HIP(hipMemcpyAsync(cell.data(), s.cell, sizeof(int32_t) * np, hipMemcpyDeviceToHost, ctx->stream));
    CHY_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < np; i++) {
      const int c = cell[i];
      const int ci[3] = {c % g.nx, (c / g.nx) % g.ny, c / (g.nx * g.ny)};
      const float o[3] = {hp4[i].x, hp4[i].y, hp4[i].z}, v[3] = {hv4[i].x, hv4[i].y, hv4[i].z};
      double r2 = 0.0, v2 = 0.0;
      for (int q = 0; q < 3; q++) {
        const double x = g.p0[q] + ((double)ci[q] + (double)o[q]) * g.d[q];
        if (pos) pos[4 * i + q] = x;
        if (vel) vel[4 * i + q] = (double)v[q];
        r2 += x * x;
        v2 += (double)v[q] * (double)v[q];
      }
      if (pos) pos[4 * i + 3] = sqrt(r2);
      if (vel) vel[4 * i + 3] = sqrt(v2);
      if (w) w[i] = ctx->weighted ? (double)hp4[i].w : 1.0;
      if (pid) memcpy(&pid[i], &hv4[i].w, sizeof(uint32_t));
    }
  }
  return CHY_OK;
}

// The reference's placements (InitializePositions, particles.c:26-45).
static int particles_init_density(chy_ctx *ctx)
{
  const chy_params &p = ctx->prm;
  const int64_t nc = ctx->g.nc;
  const int kind = p.density_type;
  static const char *const names[] = {"flat-sobol", "flat-reverse", "flat-normal", "uniform", "sinusoidal",
                                      "gaussian"};
  if (kind == CHY_DENSITY_FLAT_NORMAL || kind == CHY_DENSITY_FLAT_REVERSE || kind == CHY_DENSITY_GAUSSIAN)
    return set_err(ctx, CHY_ERR_ARG_WRONG, std::string("Not implemented: ") + names[kind] + " density");
  if (kind != CHY_DENSITY_FLAT_SOBOL && kind != CHY_DENSITY_UNIFORM && kind != CHY_DENSITY_SINUSOIDAL)
    return set_err(ctx, CHY_ERR_ARG_WRONG, "unknown density type");
  int64_t np = p.Np;
  if (kind == CHY_DENSITY_UNIFORM) {
    // Np / nc ions in each cell but the last along every axis (the
    // reference's swarm keeps the unset slots of the other cells at undefined
    // positions; here they are not created)
    np = (p.Np / nc) * (int64_t)(ctx->g.nx - 1) * (ctx->g.ny - 1) * (ctx->g.nz - 1);
  }
  // the reference's Sobol generator stops at its 2^30-th call (random.c:279-281)
  if (kind == CHY_DENSITY_FLAT_SOBOL && np >= (int64_t)1 << 30)
    return set_err(ctx, CHY_ERR_ARG_WRONG, "MAXBIT too small in the Sobol sequence: at most 2^30 - 1 ions");
  if (np > ctx->cap) return set_err(ctx, CHY_ERR_ARG_RANGE, "the ion count exceeds the context capacity");
  if (p.perturbation != CHY_PERT_NONE && !ctx->weighted)
    return set_err(ctx, CHY_ERR_ARG_INCOMP, "perturbation needs weight buffers");
  ctx->np = np;
  if (np > 0) {
    const SobolDirs sob = sobol_dirs();
    const unsigned nb = (unsigned)((np + kBlock - 1) / kBlock);
    PState<double> sd = ctx->pd[ctx->cur];
    PState<float> sf = ctx->pf[ctx->cur];
    if (ctx->mixed)
      hipLaunchKernelGGL((k_init_density<true>), nb, kBlock, 0, ctx->stream, ctx->g, kind, np, sob,
                         (uint32_t)p.seed, (uint32_t)(p.seed >> 32), p.vi0[0], p.vi0[1], p.vi0[2], p.viT[0],
                         p.viT[1], p.viT[2], p.perturbation, p.perturbation_amplitude, sd, sf);
    else
      hipLaunchKernelGGL((k_init_density<false>), nb, kBlock, 0, ctx->stream, ctx->g, kind, np, sob,
                         (uint32_t)p.seed, (uint32_t)(p.seed >> 32), p.vi0[0], p.vi0[1], p.vi0[2], p.viT[0],
                         p.viT[1], p.viT[2], p.perturbation, p.perturbation_amplitude, sd, sf);
  }
  return settle(ctx);
}

int particles_init(chy_ctx *ctx)
{
  const chy_params &p = ctx->prm;
  const int64_t nc = ctx->g.nc;
  if (p.density_type != CHY_DENSITY_LATTICE) return particles_init_density(ctx);
  int64_t ppc = p.ppc > 0 ? p.ppc : (p.Np % nc == 0 ? p.Np / nc : 0);
  if (ppc <= 0) return set_err(ctx, CHY_ERR_ARG_INCOMP, "the per-cell lattice placement needs Np to be a multiple of the cell count (-ppc)");
  const int64_t np = ppc * nc;
  if (np > ctx->cap) return set_err(ctx, CHY_ERR_ARG_RANGE, "ppc x cells exceeds the context capacity");
  if (p.perturbation != CHY_PERT_NONE && !ctx->weighted)
    return set_err(ctx, CHY_ERR_ARG_INCOMP, "perturbation needs weight buffers");
  ctx->np = np;
  const unsigned nb = (unsigned)((np + kBlock - 1) / kBlock);
  PState<double> sd = ctx->pd[ctx->cur];
  PState<float> sf = ctx->pf[ctx->cur];
  if (ctx->mixed)
    hipLaunchKernelGGL((k_init<true>), nb, kBlock, 0, ctx->stream, ctx->g, ppc, (uint32_t)p.seed,
                       (uint32_t)(p.seed >> 32), p.vi0[0], p.vi0[1], p.vi0[2], p.viT[0], p.viT[1], p.viT[2],
                       p.perturbation, p.perturbation_amplitude, sd, sf);
  else
    hipLaunchKernelGGL((k_init<false>), nb, kBlock, 0, ctx->stream, ctx->g, ppc, (uint32_t)p.seed,
                       (uint32_t)(p.seed >> 32), p.vi0[0], p.vi0[1], p.vi0[2], p.viT[0], p.viT[1], p.viT[2],
                       p.perturbation, p.perturbation_amplitude, sd, sf);
  if (ctx->weighted && p.perturbation == CHY_PERT_NONE && !ctx->mixed) launch_fill<double>(ctx, np, sd.w, 1.0);
  return settle(ctx);
}

// The push's counters, once read: collision statistics, non-finite state,
// the refusal limit.
static int push_finish(chy_ctx *ctx, const unsigned long long cs[6])
{
  ctx->push_pending = false;
  chy_collision_stats *stats = ctx->push_stats;
  ctx->push_stats = nullptr;
  if (stats) {
    stats->accepted = (int64_t)cs[0];
    stats->refused = (int64_t)cs[1];
    stats->capped = (int64_t)cs[2];
    stats->Nc = ctx->push_Nc;
    memcpy(&stats->expected, &cs[4], sizeof(double));
    memcpy(&stats->variance, &cs[5], sizeof(double));
  }
  if (cs[3]) return set_err(ctx, CHY_ERR_FP, "non-finite particle position or velocity after the push");
  const int64_t lim = ctx->prm.collision_refusal_limit;  // particles.c:603 (10)
  if (lim > 0 && (int64_t)cs[1] >= lim)
    return set_err(ctx, CHY_ERR_SIG, "Failed to collide " + std::to_string(cs[1]) + " ion-neutral pairs. Aborting.");
  return CHY_OK;
}

int particles_push_check(chy_ctx *ctx)
{
  if (!ctx->push_pending) return CHY_OK;
  const uint32_t *w[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t v[4];
  return read_status(ctx, w, v, nullptr, 0, nullptr);
}

int particles_push(chy_ctx *ctx, bool move, bool collide, chy_collision_stats *stats, bool defer)
{
  if (ctx->push_pending) TRY(particles_push_check(ctx));
  PushConst pc = push_const(ctx);
  const int64_t Nc = collide ? collision_count(ctx->prm, ctx->np) : 0;
  pc.Nc = (double)Nc;
  if (collide && Nc == 0) collide = false;
  ctx->partials_fresh = false;
  // slots 0-3 counts, 4-5 the FP64 sums of p and p (1 - p) (zero bits are 0.0)
  launch_fill<unsigned long long>(ctx, 6, ctx->cstats, 0ull);
  // the histogram is only needed by the general sort (the neighbour-move
  // sort counts per old cell)
  const bool count = move && !ctx->order_valid;
  if (count) launch_fill<uint32_t>(ctx, ctx->g.nc + 1, ctx->counts, 0u);
  // 65536 blocks (16 ions per thread at 256^3 x 16): 3.96 ms, against 4.19
  // with 4096 blocks and 4.04 with 262144
  const unsigned nb = grid_for(ctx->np, kPushBlocks);
  timer_start(ctx, "push");
#define LAUNCH(KER, T, C, M, ...)                                                                             \
  hipLaunchKernelGGL((KER<C, M>), nb, kBlock, 0, ctx->stream, ctx->g, pcT, ctx->np, st,                      \
                     (const Vec4<T> *)ctx->efield[0], ctx->S_dev, count ? ctx->counts : nullptr, ctx->cstats,   \
                     ctx->push_part, ##__VA_ARGS__)
  if (ctx->mixed) {
    const PushConstF pcT = push_const_f(pc);
    PState<float> st = ctx->pf[ctx->cur];
    // from the stable order the moving push also writes each ion's move
    // direction, which the neighbour-move sort then reads instead of the keys
    uint8_t *dir = move && ctx->order_valid && !ctx->keys_moved ? ctx->dir : nullptr;
    if (collide && move) LAUNCH(k_push_mixed, float, true, true, dir);
    else if (collide) LAUNCH(k_push_mixed, float, true, false, dir);
    else if (move) LAUNCH(k_push_mixed, float, false, true, dir);
    else LAUNCH(k_push_mixed, float, false, false, dir);
    if (move) ctx->dir_valid = dir != nullptr;
  } else {
    const PushConst &pcT = pc;
    PState<double> st = ctx->pd[ctx->cur];
    uint8_t *dir = move && ctx->order_valid && !ctx->keys_moved ? ctx->dir : nullptr;
    if (collide && move) LAUNCH(k_push_fp64, double, true, true, dir);
    else if (collide) LAUNCH(k_push_fp64, double, true, false, dir);
    else if (move) LAUNCH(k_push_fp64, double, false, true, dir);
    else LAUNCH(k_push_fp64, double, false, false, dir);
    if (move) ctx->dir_valid = dir != nullptr;
  }
#undef LAUNCH
  if (collide && ctx->mixed) {  // the blocks' tallies, summed in a fixed order
    double *tmp = ctx->push_part + 4 * (int64_t)kPushBlocks;
    hipLaunchKernelGGL((k_push_reduce<false>), kPushRed, 256, 0, ctx->stream, (const double *)ctx->push_part, (int)nb,
                       tmp, ctx->cstats);
    hipLaunchKernelGGL((k_push_reduce<true>), 1, 256, 0, ctx->stream, (const double *)tmp, kPushRed, (double *)nullptr,
                       ctx->cstats);
  }
  if (move) ctx->keys_moved = true;
  timer_stop(ctx, "push");
  // the counters are read now, or (defer) with the next status read-back,
  // the sort's, and checked before its own results
  ctx->push_pending = true;
  ctx->push_Nc = Nc;
  ctx->push_stats = stats;
  return defer ? CHY_OK : particles_push_check(ctx);
}

int particles_move_only(chy_ctx *ctx)
{
  PushConst pc = push_const(ctx);
  ctx->dir_valid = false;  // keys change without directions
  ctx->keys_moved = true;
  ctx->partials_fresh = false;
  launch_fill<unsigned long long>(ctx, 4, ctx->cstats, 0ull);
  launch_fill<uint32_t>(ctx, ctx->g.nc + 1, ctx->counts, 0u);
  const unsigned nb = grid_for(ctx->np);
  if (ctx->mixed)
    hipLaunchKernelGGL(k_move_mixed, nb, kBlock, 0, ctx->stream, ctx->g, push_const_f(pc), ctx->np, ctx->pf[ctx->cur], ctx->counts,
                       ctx->cstats);
  else
    hipLaunchKernelGGL(k_move_fp64, nb, kBlock, 0, ctx->stream, ctx->g, pc, ctx->np, ctx->pd[ctx->cur], ctx->counts,
                       ctx->cstats);
  unsigned long long cs[4];
  {
    const uint32_t *w[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t v[4];
    TRY(read_status(ctx, w, v, ctx->cstats, 4, cs));
  }
  if (cs[3]) return set_err(ctx, CHY_ERR_FP, "non-finite particle position after the move");
  return CHY_OK;
}

}  // namespace chy
