// Ray marching through an RSDF isosurface and the mmWave channel it returns:
// the consumer of the RSDF (ray_tracing_kernel.cu:60-421 and
// isosurface_optimization.py:85-246).
//
// Per antenna, width x height rays leave the antenna through a pinhole
// camera looking along -rx_dir (ray_tracing_kernel.cu:60-127); each ray is
// sphere-traced through the trilinearly interpolated SDF grid minus the
// isosurface offset tau, stepping by the SDF value and taking MIN_STEP where
// a corner is undefined, until the value drops below one voxel
// (ray_tracing_kernel.cu:206-421).  The rough hit is refined by one step of
// the interpolated value along the ray (isosurface_optimization.py:85-109),
// and the channel is the sum over hits of exp(+j 2 pi f d / c) with d twice
// the antenna-to-hit distance (+0.15 m for the TI radar)
// (isosurface_optimization.py:203-208).
//
// Arithmetic follows the reference expression by expression (FP32, no FMA
// contraction; the pixel angle in FP64 as its mixed-precision expression
// evaluates; distances in FP64 from FP32 hits), except where noted:
//   * grid reads clamp their indices into the grid (the reference reads out
//     of bounds for a hit on the far faces);
//   * the channel sum uses the FP64-anchored phase recurrence of the
//     backward (anchor every K = 32 samples) instead of one FP64 exp per
//     (ray, sample), and FP32 lane sums reduced in FP64 in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "mmn_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxSteps = 1000;        // ray_tracing_kernel.cu:342
constexpr float kUndef = -50.0f;       // UNDEF_VAL
constexpr float kMinStep = 0.005f;     // MIN_STEP
constexpr int kK = 32;                 // samples per wave of the synthesis
constexpr int kMaxSims = 65535;        // simulations of one batched launch (grid y)

struct RayGeom {
  float fov_w, fov_h;       // degrees
  float bmin[3], bmax[3];   // grid bounding box (voxel centres of the corners)
  int res[3];
  float eye[3], rxd[3];
  int width, height;
  float tau;                // isosurface offset subtracted from the grid
  int small;                // NX NY NZ <= 2^24: corner indices in integers (value_at)
};

// One simulation of a batch (mmn_simulate_channels): the antenna, its look
// direction, the isosurface offset, the grid component and the receiver
// position of the FP64 round trip.  64 bytes, read with scalar loads.
struct SimParam {
  float eye[3];
  float rxd[3];
  float tau;
  int32_t comp;
  double rx[3];
  double pad;
};
static_assert(sizeof(SimParam) == 64, "SimParam is one 64-byte record");

__device__ __forceinline__ float len3(float x, float y, float z) {
#pragma clang fp contract(off)
  return sqrtf(x * x + y * y + z * z);
}

// GenerateRay (ray_tracing_kernel.cu:60-127): the unit direction of ray p.
__device__ void ray_dir(const RayGeom &g, int p, float &dx, float &dy, float &dz) {
#pragma clang fp contract(off)
  const float pi = 3.14159265358979323846f;
  const float vfov = g.fov_h * (pi / 180.f);
  const float hfov = g.fov_w * (pi / 180.f);
  const float lw = len3(g.rxd[0], g.rxd[1], g.rxd[2]);
  const float wx = g.rxd[0] / lw, wy = g.rxd[1] / lw, wz = g.rxd[2] / lw;
  // up = (1, 0, 0): cross(up, w) = (0, -wz, wy)
  const float cx = 0.0f * wz - 0.0f * wy, cy = 0.0f * wx - 1.0f * wz, cz = 1.0f * wy - 0.0f * wx;
  const float lc = len3(cx, cy, cz);
  const float ux = cx / lc, uy = cy / lc, uz = cz / lc;
  const float vx = wy * uz - wz * uy, vy = wz * ux - wx * uz, vz = wx * uy - wy * ux;
  const int x = p % g.width, y = p / g.width;
  // (float)x + 0.5 is a double expression in the reference
  const float ha = (float)(-(double)hfov + (double)(hfov * 2) * ((double)(float)x + 0.5) /
                                               (double)(float)g.width);
  const float va = (float)(-(double)vfov + (double)(vfov * 2) * ((double)(float)y + 0.5) /
                                               (double)(float)g.height);
  const float px = tanf(ha), py = tanf(va);
  const float sx = px * ux + py * vx - wx;
  const float sy = px * uy + py * vy - wy;
  const float sz = px * uz + py * vz - wz;
  const float ls = len3(sx, sy, sz);
  dx = sx / ls;
  dy = sy / ls;
  dz = sz / ls;
}

__device__ __forceinline__ bool inside(const RayGeom &g, float x, float y, float z) {
  return x >= g.bmin[0] && x <= g.bmax[0] && y >= g.bmin[1] && y <= g.bmax[1] &&
         z >= g.bmin[2] && z <= g.bmax[2];
}

// Distance (ray_tracing_kernel.cu:146-196): along the ray to the box, from
// outside; -1 if no face is ahead.  The direction is renormalised one
// component at a time, each with the already updated components.
__device__ float box_distance(const RayGeom &g, float px, float py, float pz, float dx,
                              float dy, float dz) {
#pragma clang fp contract(off)
  float dist = -1.f;
  dx = dx / len3(dx, dy, dz);
  dy = dy / len3(dx, dy, dz);
  dz = dz / len3(dx, dy, dz);
  const float v[3] = {px, py, pz}, d[3] = {dx, dy, dz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (dist == -1) {
      if ((v[a] < g.bmin[a]) && (d[a] > 0)) dist = (g.bmin[a] - v[a]) / d[a];
      if ((v[a] > g.bmax[a]) && (d[a] < 0)) dist = (g.bmax[a] - v[a]) / d[a];
    } else {
      if ((v[a] < g.bmin[a]) && (d[a] > 0)) dist = fmaxf(dist, (g.bmin[a] - v[a]) / d[a]);
      if ((v[a] > g.bmax[a]) && (d[a] < 0)) dist = fmaxf(dist, (g.bmax[a] - v[a]) / d[a]);
    }
  }
  return dist;
}

// flat() of the reference: the index is formed in FP32 from truncated
// coordinates (exact below 2^24 voxels); clamped into the grid here.
__device__ __forceinline__ int64_t flat_idx(const RayGeom &g, float x, float y, float z) {
#pragma clang fp contract(off)
  const float fz = (float)(int)z, fy = (float)(int)y, fx = (float)(int)x;
  const float f = fz + fy * (float)g.res[2] + fx * (float)g.res[2] * (float)g.res[1];
  int64_t i = (int64_t)(int)f;
  const int64_t V = (int64_t)g.res[0] * g.res[1] * g.res[2];
  return i < 0 ? 0 : (i >= V ? V - 1 : i);
}

// The cell of ValueAt (ray_tracing_kernel.cu:255-270) for a point inside the
// box: its minimum corner as the reference's float indices, and the grid at
// its eight corners (before tau is subtracted), in the order
// 000, 010, 001, 011, 100, 110, 101, 111 (x, y, z offsets).
__device__ __forceinline__ void cell_at(const float *grid, const RayGeom &g, float px, float py,
                                        float pz, float &ix, float &iy, float &iz,
                                        float (&c)[8]) {
#pragma clang fp contract(off)
  const float vs = (g.bmax[1] - g.bmin[1]) / (float)(g.res[1] - 1);
  ix = floorf((px - g.bmin[0]) / vs);
  iy = floorf((py - g.bmin[1]) / vs);
  iz = floorf((pz - g.bmin[2]) / vs);
  if (ix == (g.bmax[0] - g.bmin[0]) / vs) ix = (g.bmax[0] - g.bmin[0]) / vs - 1;
  if (iy == (g.bmax[1] - g.bmin[1]) / vs) iy = (g.bmax[1] - g.bmin[1]) / vs - 1;
  if (iz == (g.bmax[2] - g.bmin[2]) / vs) iz = (g.bmax[2] - g.bmin[2]) / vs - 1;
  if (g.small) {
    // flat_idx sums non-negative integer terms in FP32: below 2^24 every
    // partial sum is exact, and a true index at or above 2^24 (>= V here)
    // stays there and is clamped to V - 1 either way.  So the corners are the
    // base index plus their offsets in integers, clamped the same.
    const int64_t s2 = g.res[2], s1 = (int64_t)g.res[1] * g.res[2], V = g.res[0] * s1;
    const int64_t b = (int64_t)(int)iz + (int64_t)(int)iy * s2 + (int64_t)(int)ix * s1;
    auto at = [&](int64_t o) { return grid[min(b + o, V - 1)]; };
    c[0] = at(0), c[1] = at(s2), c[2] = at(1), c[3] = at(s2 + 1);
    c[4] = at(s1), c[5] = at(s1 + s2), c[6] = at(s1 + 1), c[7] = at(s1 + s2 + 1);
  } else {
    c[0] = grid[flat_idx(g, ix, iy, iz)], c[1] = grid[flat_idx(g, ix, iy + 1, iz)];
    c[2] = grid[flat_idx(g, ix, iy, iz + 1)], c[3] = grid[flat_idx(g, ix, iy + 1, iz + 1)];
    c[4] = grid[flat_idx(g, ix + 1, iy, iz)], c[5] = grid[flat_idx(g, ix + 1, iy + 1, iz)];
    c[6] = grid[flat_idx(g, ix + 1, iy, iz + 1)];
    c[7] = grid[flat_idx(g, ix + 1, iy + 1, iz + 1)];
  }
}

// A corner is undefined when grid - tau < UNDEF_VAL + 10 (ray_tracing_kernel.cu:274-285).
__device__ __forceinline__ bool any_undefined(const float (&c)[8], float tau) {
#pragma clang fp contract(off)
  const float lim = kUndef + 10;
  bool u = false;
#pragma unroll
  for (int k = 0; k < 8; ++k) u |= c[k] - tau < lim;
  return u;
}

// ValueAt (ray_tracing_kernel.cu:206-301) on grid - tau.
__device__ float value_at(const float *grid, const RayGeom &g, float px, float py, float pz,
                          float dx, float dy, float dz, bool first) {
#pragma clang fp contract(off)
  if (!inside(g, px, py, pz)) {
    if (first) return box_distance(g, px, py, pz, dx, dy, dz) + 0.00001f;
    return -1;
  }
  const float vs = (g.bmax[1] - g.bmin[1]) / (float)(g.res[1] - 1);
  float ix, iy, iz, c[8];
  cell_at(grid, g, px, py, pz, ix, iy, iz, c);
  if (any_undefined(c, g.tau)) return kMinStep;
  const float g000 = c[0] - g.tau, g010 = c[1] - g.tau, g001 = c[2] - g.tau;
  const float g011 = c[3] - g.tau, g100 = c[4] - g.tau, g110 = c[5] - g.tau;
  const float g101 = c[6] - g.tau, g111 = c[7] - g.tau;
  const float tx = (px - (g.bmin[0] + ix * vs)) / vs;
  const float c01 = (1.f - tx) * g000 + tx * g100;
  const float c23 = (1.f - tx) * g010 + tx * g110;
  const float c45 = (1.f - tx) * g001 + tx * g101;
  const float c67 = (1.f - tx) * g011 + tx * g111;
  const float ty = (py - (g.bmin[1] + iy * vs)) / vs;
  const float c0 = (1.f - ty) * c01 + ty * c23;
  const float c1 = (1.f - ty) * c45 + ty * c67;
  const float tz = (pz - (g.bmin[2] + iz * vs)) / vs;
  return (1.f - tz) * c0 + tz * c1;
}

// The rough hit at the march point r and the voxel index of its minimum
// corner (ray_tracing_kernel.cu:383-400).
__device__ __forceinline__ void hit_at(const RayGeom &g, const float (&r)[3], float (&hit)[3],
                                       float (&vox)[3]) {
#pragma clang fp contract(off)
  const float vs = (g.bmax[1] - g.bmin[1]) / (float)(g.res[1] - 1);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float v = floorf((r[a] - g.bmin[a]) / vs);
    if (v == (float)(g.res[a] - 1)) v = v - 1;
    vox[a] = v;
    hit[a] = r[a];
  }
}

// Intersect (ray_tracing_kernel.cu:305-421) for one ray from step s0 at
// depth `depth` (0 and 0.0 for the whole march; the first step is the one
// that may enter the box): the rough hit and the voxel index of its minimum
// corner, or all -1.
__device__ void intersect_from(const float *grid, const RayGeom &g, float dx, float dy, float dz,
                               float depth, int s0, float (&hit)[3], float (&vox)[3]) {
#pragma clang fp contract(off)
  const float vs = (g.bmax[1] - g.bmin[1]) / (float)(g.res[1] - 1);
  for (int s = s0; s < kMaxSteps; ++s) {
    const float r[3] = {g.eye[0] + depth * dx, g.eye[1] + depth * dy, g.eye[2] + depth * dz};
    const float dist = value_at(grid, g, r[0], r[1], r[2], dx, dy, dz, s == 0);
    if (dist == -1) break;
    if (dist < vs && dist != 0.0f) {
      hit_at(g, r, hit, vox);
      return;
    }
    depth += dist;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) hit[a] = vox[a] = -1.0f;
}

__device__ __forceinline__ void intersect(const float *grid, const RayGeom &g, float dx,
                                          float dy, float dz, float (&hit)[3],
                                          float (&vox)[3]) {
  intersect_from(grid, g, dx, dy, dz, 0.0f, 0, hit, vox);
}

// ray_matching drop-in: rays, march, outputs as the reference's three
// (width, height, 3) tensors in pixel order.
__global__ __launch_bounds__(kBlock) void k_ray_match(float *__restrict__ hit_out,
                                                      float *__restrict__ vox_out,
                                                      float *__restrict__ dir_out,
                                                      const float *__restrict__ grid,
                                                      RayGeom g) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= g.width * g.height) return;
  float dx, dy, dz, hit[3], vox[3];
  ray_dir(g, p, dx, dy, dz);
  intersect(grid, g, dx, dy, dz, hit, vox);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    hit_out[3 * (size_t)p + a] = hit[a];
    vox_out[3 * (size_t)p + a] = vox[a];
  }
  dir_out[3 * (size_t)p + 0] = dx;
  dir_out[3 * (size_t)p + 1] = dy;
  dir_out[3 * (size_t)p + 2] = dz;
}

// Grid index of the refine step as torch indexing reads it: -1 was replaced
// by 0 (isosurface_optimization.py:185-187); other negatives wrap.
__device__ __forceinline__ int64_t torch_index(float v, int res) {
  int64_t i = (int64_t)v;
  if (i == -1) i = 0;
  if (i < 0) i += res;
  return i < 0 ? 0 : (i >= res ? res - 1 : i);
}

// The refine step (isosurface_optimization.py:85-109,190-196) of a ray that
// hit and its round trip in FP64 (isosurface_optimization.py:198-201).
__device__ double refine_dist(const float *grid, const RayGeom &g, float vsz,
                              const double (&rx)[3], int is_ti, const float (&d)[3],
                              const float (&hit)[3], const float (&vox)[3]) {
#pragma clang fp contract(off)
  int64_t ii[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) ii[a] = torch_index(vox[a], g.res[a]);
  const int64_t ny = g.res[1], nz = g.res[2];
  auto at = [&](int64_t x, int64_t y, int64_t z) {
    x = x >= g.res[0] ? g.res[0] - 1 : x;
    y = y >= g.res[1] ? g.res[1] - 1 : y;
    z = z >= g.res[2] ? g.res[2] - 1 : z;
    return grid[(x * ny + y) * nz + z] - g.tau;
  };
  float t[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float vmin = g.bmin[a] + vox[a] * vsz;  // bounding_box_lower + index * voxel_size
    t[a] = (hit[a] - vmin) / vsz;
  }
  const int64_t x = ii[0], y = ii[1], z = ii[2];
  const float c01 = (1 - t[0]) * at(x, y, z) + t[0] * at(x + 1, y, z);
  const float c23 = (1 - t[0]) * at(x, y + 1, z) + t[0] * at(x + 1, y + 1, z);
  const float c45 = (1 - t[0]) * at(x, y, z + 1) + t[0] * at(x + 1, y, z + 1);
  const float c67 = (1 - t[0]) * at(x, y + 1, z + 1) + t[0] * at(x + 1, y + 1, z + 1);
  const float c0 = (1 - t[1]) * c01 + t[1] * c23;
  const float c1 = (1 - t[1]) * c45 + t[1] * c67;
  const float sv = (1 - t[2]) * c0 + t[2] * c1;
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float q = hit[a] + d[a] * sv;  // refined hit (mask = 1)
    const double e = rx[a] - (double)q;
    s += e * e;
  }
  double dd = sqrt(s) * 2;
  if (is_ti) dd += 0.15;
  return dd;
}

// Simulation blockIdx.y of a batch: its antenna, direction, offset and
// receiver; returns the offset of its component in the grid.
__device__ __forceinline__ int64_t apply_sim(const SimParam &q, RayGeom &g, double (&rx)[3],
                                             int64_t comp_stride) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    g.eye[a] = q.eye[a];
    g.rxd[a] = q.rxd[a];
    rx[a] = q.rx[a];
  }
  g.tau = q.tau;
  return (int64_t)q.comp * comp_stride;
}

// Per ray: march, refine and the FP64 round trip; NaN marks a ray without a
// hit.  BATCH: simulation s = blockIdx.y takes its values from sp[s]
// (apply_sim) and writes dist[s * rays + p]; a ray's arithmetic is the same
// either way.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_ray_dist(double *__restrict__ dist,
                                                     const float *__restrict__ grid, RayGeom g,
                                                     float vsz, double rx0, double rx1,
                                                     double rx2, int is_ti,
                                                     const SimParam *__restrict__ sp,
                                                     int64_t comp_stride) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= g.width * g.height) return;
  double rx[3] = {rx0, rx1, rx2};
  if constexpr (BATCH) {
    grid += apply_sim(sp[blockIdx.y], g, rx, comp_stride);
    dist += (int64_t)blockIdx.y * g.width * g.height;
  }
  float d[3], hit[3], vox[3];
  ray_dir(g, p, d[0], d[1], d[2]);
  intersect(grid, g, d[0], d[1], d[2], hit, vox);
  // mask = voxel index x != -1
  dist[p] = vox[0] == -1.0f ? (double)NAN : refine_dist(grid, g, vsz, rx, is_ti, d, hit, vox);
}

// ---------------------------------------------------------------------------
// The offset-independent part of the march.  While a ray is outside the box,
// or in a cell with an undefined corner, its step does not depend on the
// isosurface offset tau: whether grid - tau < UNDEF_VAL + 10 at some corner
// only turns from false to true as tau grows (the rounded difference falls
// monotonically), so a cell undefined at the smallest offset tau_lo of a
// search is undefined at every offset >= tau_lo, and its step is MIN_STEP.
// k_ray_prefix marches each ray until its first step in a cell that is fully
// defined at tau_lo and records that step and depth (RESUME), a hit reached
// before it (HIT: the first step's box-distance quirk, or MIN_STEP below a
// voxel), or the end of the march without a hit (NONE).  Resuming the march
// from the recorded step for any tau >= tau_lo repeats exactly what the whole
// march computes, and a NONE ray has no hit at any such tau.
// ---------------------------------------------------------------------------
constexpr int kStNone = 0, kStResume = 1, kStHit = 2;

struct RayState {
  float depth;
  int code;  // status | step << 2
};

__device__ int march_prefix(const float *grid, const RayGeom &g, float dx, float dy, float dz,
                            float tau_lo, float &depth) {
#pragma clang fp contract(off)
  const float vs = (g.bmax[1] - g.bmin[1]) / (float)(g.res[1] - 1);
  depth = 0.0f;
  for (int s = 0; s < kMaxSteps; ++s) {
    const float r[3] = {g.eye[0] + depth * dx, g.eye[1] + depth * dy, g.eye[2] + depth * dz};
    float dist;
    if (!inside(g, r[0], r[1], r[2])) {
      dist = s == 0 ? box_distance(g, r[0], r[1], r[2], dx, dy, dz) + 0.00001f : -1.0f;
    } else {
      float ix, iy, iz, c[8];
      cell_at(grid, g, r[0], r[1], r[2], ix, iy, iz, c);
      if (!any_undefined(c, tau_lo)) return kStResume | (s << 2);
      dist = kMinStep;
    }
    if (dist == -1) return kStNone;
    if (dist < vs && dist != 0.0f) return kStHit | (s << 2);
    depth += dist;
  }
  return kStNone;
}

__global__ __launch_bounds__(kBlock) void k_ray_prefix(RayState *__restrict__ st,
                                                       const float *__restrict__ grid,
                                                       RayGeom g, const SimParam *__restrict__ sp,
                                                       int64_t comp_stride, float tau_lo) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  const int n = g.width * g.height;
  if (p >= n) return;
  double rx[3];
  grid += apply_sim(sp[blockIdx.y], g, rx, comp_stride);
  float d[3], depth;
  ray_dir(g, p, d[0], d[1], d[2]);
  const int code = march_prefix(grid, g, d[0], d[1], d[2], tau_lo, depth);
  st[(int64_t)blockIdx.y * n + p] = RayState{depth, code};
}

// The live rays (status RESUME or HIT) of each class q < Q, the contiguous
// rays [q * per, (q + 1) * per) with per = ceil(n / Q): counted, scanned into
// offsets off[s][q] (off[s][Q] = the simulation's live rays) and listed, so
// that each simulation's list holds its live rays in ray order.
__device__ __forceinline__ void class_range(int q, int n, int Q, int &r, int &r1) {
  const int64_t per = ((int64_t)n + Q - 1) / Q;
  r = (int)min((int64_t)n, q * per);
  r1 = (int)min((int64_t)n, r + per);
}

__global__ __launch_bounds__(kBlock) void k_live_count(int *__restrict__ off,
                                                       const RayState *__restrict__ st, int n,
                                                       int Q) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= Q) return;
  const RayState *sts = st + (int64_t)blockIdx.y * n;
  int r, r1, k = 0;
  class_range(q, n, Q, r, r1);
  for (; r < r1; ++r) k += (sts[r].code & 3) != kStNone;
  off[(int64_t)blockIdx.y * (Q + 1) + q] = k;
}

// Exclusive scan of one simulation's Q counts (one workgroup of 1024 threads
// per simulation, each thread a contiguous run); the total goes to off[Q]
// and to tot[s].
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_live_scan(int *__restrict__ off,
                                                            int *__restrict__ tot, int Q) {
  __shared__ int part[kScanThreads];
  int *o = off + (int64_t)blockIdx.x * (Q + 1);
  const int per = (Q + kScanThreads - 1) / kScanThreads;
  const int lo = min(Q, (int)threadIdx.x * per), hi = min(Q, lo + per);
  int sum = 0;
  for (int q = lo; q < hi; ++q) sum += o[q];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int d = 1; d < kScanThreads; d <<= 1) {
    const int v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - sum;  // exclusive prefix of this run
  for (int q = lo; q < hi; ++q) {
    const int v = o[q];
    o[q] = run;
    run += v;
  }
  if (threadIdx.x == kScanThreads - 1) {
    o[Q] = part[kScanThreads - 1];
    tot[blockIdx.x] = part[kScanThreads - 1];
  }
}

__global__ __launch_bounds__(kBlock) void k_live_fill(int *__restrict__ list,
                                                      const int *__restrict__ off,
                                                      const RayState *__restrict__ st, int n,
                                                      int Q) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= Q) return;
  const RayState *sts = st + (int64_t)blockIdx.y * n;
  int *ls = list + (int64_t)blockIdx.y * n;
  int r, r1, k = off[(int64_t)blockIdx.y * (Q + 1) + q];
  class_range(q, n, Q, r, r1);
  for (; r < r1; ++r)
    if ((sts[r].code & 3) != kStNone) ls[k++] = r;
}

// One offset of a search from the recorded prefix: each live ray t (of
// off[s][Q]) resumes its march (or takes its recorded hit) with the
// simulation's tau, refines and writes its round trip (NaN: no hit) to
// dl[s * live_max + t].
__global__ __launch_bounds__(kBlock) void k_ray_resume(double *__restrict__ dl,
                                                       const float *__restrict__ grid,
                                                       RayGeom g, float vsz, int is_ti,
                                                       const SimParam *__restrict__ sp,
                                                       int64_t comp_stride,
                                                       const RayState *__restrict__ st,
                                                       const int *__restrict__ list,
                                                       const int *__restrict__ off, int Q,
                                                       int live_max) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int n = g.width * g.height;
  const int live = min(off[(int64_t)blockIdx.y * (Q + 1) + Q], live_max);
  if (t >= live) return;
  double rx[3];
  grid += apply_sim(sp[blockIdx.y], g, rx, comp_stride);
  const int r = min(max(list[(int64_t)blockIdx.y * n + t], 0), n - 1);
  const RayState q = st[(int64_t)blockIdx.y * n + r];
  float d[3], hit[3], vox[3];
  ray_dir(g, r, d[0], d[1], d[2]);
  if ((q.code & 3) == kStHit) {
    const float p3[3] = {g.eye[0] + q.depth * d[0], g.eye[1] + q.depth * d[1],
                         g.eye[2] + q.depth * d[2]};
    hit_at(g, p3, hit, vox);
  } else {
    intersect_from(grid, g, d[0], d[1], d[2], q.depth, q.code >> 2, hit, vox);
  }
  dl[(int64_t)blockIdx.y * live_max + t] =
      vox[0] == -1.0f ? (double)NAN : refine_dist(grid, g, vsz, rx, is_ti, d, hit, vox);
}

// Channel synthesis: out[f] = sum over hit rays of exp(+j 2 pi (f0 + f df) d / c).
//
// The round trips of a simulation (NaN for a ray without a hit) are first
// compacted, in order, into its dense hit list (k_hit_count, k_live_scan,
// k_hit_fill).  The list is split into nc(H) near-equal chunks of
// consecutive hits (H hits; at least kHitsPerChunk a chunk, at most
// kMaxChunks chunks); one wave per (chunk, block of kK samples) walks its
// chunk's hits, lane l taking hits l, l + 64, ..., and keeps kK complex sums
// per lane from one FP64-reduced anchor and the two-term recurrence
// z_{t+1} = 2 cos(w) z_t - z_{t-1} per hit; a butterfly reduces the wave, and
// the chunks' partials are added in chunk order in FP64 (k_synth_reduce).
// Every step depends only on the hit list, so any two paths that produce the
// same hits in the same order (the full march, the batched march, the
// prefixed march of the live rays) give the same channel bit for bit.
constexpr int kHitsPerChunk = 512;
constexpr int kMaxChunks = 64;
constexpr int kCompactPer = 4;                    // entries per thread of a tile
constexpr int kCompactTile = kBlock * kCompactPer;

__device__ __host__ __forceinline__ int hit_chunks(int hits) {
  const int c = (hits + kHitsPerChunk - 1) / kHitsPerChunk;
  return c < kMaxChunks ? c : kMaxChunks;
}

// One hit's kK terms added to the lane's sums.
__device__ __forceinline__ void add_ray(double d, double fc, double dfc, float (&accr)[kK],
                                        float (&acci)[kK]) {
#pragma clang fp contract(off)
  float ar, ai, rr, ri;
  phasor_hw(d * fc, ar, ai);
  phasor_hw(d * dfc, rr, ri);
  float zr = ar, zi = -ai;          // exp(+j ...)
  const float sr = rr, si = -ri;
  float pr = zr * sr - zi * si, pi_ = zr * si + zi * sr;
  const float c2 = sr + sr;
#pragma unroll
  for (int t = 0; t < kK; ++t) {
    accr[t] += zr;
    acci[t] += zi;
    const float nr = fmaf(c2, pr, -zr), ni = fmaf(c2, pi_, -zi);
    zr = pr;
    zi = pi_;
    pr = nr;
    pi_ = ni;
  }
}

// The wave's sums (butterfly), sample t to lane t.
__device__ __forceinline__ void store_wave(float2 *__restrict__ out, int lane, int nf,
                                           float (&accr)[kK], float (&acci)[kK]) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
#pragma unroll
    for (int t = 0; t < kK; ++t) {
      accr[t] += __shfl_xor(accr[t], o, kWave);
      acci[t] += __shfl_xor(acci[t], o, kWave);
    }
  }
#pragma unroll
  for (int t = 0; t < kK; ++t)
    if (lane == t && t < nf) out[t] = make_float2(accr[t], acci[t]);
}

// Source length of simulation s: its live-ray count from the prefix's
// offsets (capped at the buffer's stride), or the fixed ray count.
__device__ __forceinline__ int64_t src_len(const int *len, int64_t len_stride, int64_t m,
                                           int s) {
  return len ? min((int64_t)len[(int64_t)s * len_stride], m) : m;
}

// Tile b of simulation s (kCompactTile entries, entry e * kBlock + t of the
// tile to thread t): which entries hold a hit, and each hit's rank among the
// tile's hits (wave ballots; wave totals in LDS).  Returns the tile's count.
__device__ __forceinline__ int tile_hits(const double *__restrict__ src, int64_t base,
                                         int64_t m, double (&v)[kCompactPer],
                                         int (&rank)[kCompactPer]) {
  __shared__ int wtot[kCompactPer * kWavesPerBlock];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  bool ok[kCompactPer];
#pragma unroll
  for (int e = 0; e < kCompactPer; ++e) {
    const int64_t p = base + (int64_t)e * kBlock + threadIdx.x;
    v[e] = p < m ? src[p] : (double)NAN;
    ok[e] = v[e] == v[e];
  }
#pragma unroll
  for (int e = 0; e < kCompactPer; ++e) {
    const uint64_t bal = __ballot(ok[e]);
    rank[e] = ok[e] ? __popcll(bal & below) : -1;
    if (lane == 0) wtot[e * kWavesPerBlock + w] = __popcll(bal);
  }
  __syncthreads();
  // exclusive prefix over (e, wave) in entry order: e-major, then wave
  int pre[kCompactPer], run = 0;
#pragma unroll
  for (int k = 0; k < kCompactPer * kWavesPerBlock; ++k) {
    const int c = wtot[k];
#pragma unroll
    for (int e = 0; e < kCompactPer; ++e)
      if (k == e * kWavesPerBlock + w) pre[e] = run;
    run += c;
  }
  const int total = run;
#pragma unroll
  for (int e = 0; e < kCompactPer; ++e)
    if (rank[e] >= 0) rank[e] += pre[e];
  return total;
}

__global__ __launch_bounds__(kBlock) void k_hit_count(int *__restrict__ cnt,
                                                      const double *__restrict__ src,
                                                      int64_t stride, const int *len,
                                                      int64_t len_stride, int nb) {
  const int s = blockIdx.y;
  const int64_t m = src_len(len, len_stride, stride, s);
  const int64_t base = (int64_t)blockIdx.x * kCompactTile;
  double v[kCompactPer];
  int rank[kCompactPer];
  const int total = base < m ? tile_hits(src + (int64_t)s * stride, base, m, v, rank) : 0;
  if (threadIdx.x == 0) cnt[(int64_t)s * (nb + 1) + blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void k_hit_fill(double *__restrict__ dense,
                                                     const int *__restrict__ off,
                                                     const double *__restrict__ src,
                                                     int64_t stride, const int *len,
                                                     int64_t len_stride, int nb) {
  const int s = blockIdx.y;
  const int64_t m = src_len(len, len_stride, stride, s);
  const int64_t base = (int64_t)blockIdx.x * kCompactTile;
  if (base >= m) return;
  double v[kCompactPer];
  int rank[kCompactPer];
  tile_hits(src + (int64_t)s * stride, base, m, v, rank);
  double *out = dense + (int64_t)s * stride + off[(int64_t)s * (nb + 1) + blockIdx.x];
#pragma unroll
  for (int e = 0; e < kCompactPer; ++e)
    if (rank[e] >= 0) out[rank[e]] = v[e];
}

// Wave (chunk c, sample block fb) of simulation blockIdx.y; hits[s] is the
// simulation's hit count (off[s][nb] of the compaction).
__global__ __launch_bounds__(kBlock) void k_synth_dense(float2 *__restrict__ part,
                                                        const double *__restrict__ dense,
                                                        int64_t stride,
                                                        const int *__restrict__ hits,
                                                        int64_t hits_stride, double f0c,
                                                        double dfc, int F, int nfb) {
  const int s = blockIdx.y;
  const int lane = threadIdx.x % kWave;
  const int item = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  const int c = item / nfb, fb = item % nfb;
  const int H = hits[(int64_t)s * hits_stride];
  const int nc = hit_chunks(H);
  if (c >= nc) return;
  const int per = (H + nc - 1) / nc;
  const int t0 = c * per, t1 = min(H, t0 + per);
  const double *d = dense + (int64_t)s * stride;
  const int f_lo = fb * kK;
  const double fc = fma((double)f_lo, dfc, f0c);
  float accr[kK], acci[kK];
#pragma unroll
  for (int t = 0; t < kK; ++t) accr[t] = acci[t] = 0.0f;
  for (int t = t0 + lane; t < t1; t += kWave) add_ray(d[t], fc, dfc, accr, acci);
  store_wave(part + ((int64_t)s * kMaxChunks + c) * F + f_lo, lane, min(kK, F - f_lo), accr,
             acci);
}

// The chunks' partials of each sample added in chunk order in FP64 (16 loads
// issued before their additions, which keep their order); a workgroup is one
// wave, so a few hundred samples spread over as many CUs.  No hit: zeros.
constexpr int kReduceBlock = 64;
__global__ __launch_bounds__(kReduceBlock) void k_synth_reduce(float2 *__restrict__ out,
                                                               const float2 *__restrict__ part,
                                                               const int *__restrict__ hits,
                                                               int64_t hits_stride, int F) {
  const int f = blockIdx.x * kReduceBlock + threadIdx.x;
  if (f >= F) return;
  const int nchunks = hit_chunks(hits[(int64_t)blockIdx.y * hits_stride]);
  out += (int64_t)blockIdx.y * F;
  part += (int64_t)blockIdx.y * kMaxChunks * F + f;
  double sr = 0.0, si = 0.0;
  int c = 0;
  for (; c + 16 <= nchunks; c += 16) {
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = part[(size_t)(c + k) * F];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      sr += v[k].x;
      si += v[k].y;
    }
  }
  for (; c < nchunks; ++c) {
    const float2 v = part[(size_t)c * F];
    sr += v.x;
    si += v.y;
  }
  out[f] = make_float2((float)sr, (float)si);
}

// Range-profile loss of isosurface_optimization.py:211-246 for B channel
// pairs: DFT of each channel zero-padded to n_fft, bins [lo, hi) (target
// bins >= hi_target zeroed first), each profile normalised to unit L2 norm,
// loss = sum | |sim| - |target| |.  One block per pair; twiddles from the
// exact integer phase k n mod n_fft, tabulated once per block in LDS when
// n_fft <= kLossTable (the same sincospi values), and the pair's samples
// staged in LDS when F <= kLossStage.
constexpr int kLossTable = 2048;
constexpr int kLossStage = 1024;
__global__ __launch_bounds__(kBlock) void k_channel_loss(float *__restrict__ loss,
                                                         const float2 *__restrict__ sim,
                                                         const float2 *__restrict__ tgt,
                                                         int F, int n_fft, int lo, int hi,
                                                         int hi_target) {
  // No contraction: the simulated and the target channel go through the same
  // rounded operations, so a channel against itself gives a loss of exactly
  // 0, as the reference's difference of two normalised tensors does (a fused
  // multiply-add can contract the two sides differently).
#pragma clang fp contract(off)
  __shared__ double sh[3][kBlock];
  __shared__ float mag[2][1024];
  __shared__ double2 tw[kLossTable];        // (cos, sin)(2 pi m / n_fft)
  __shared__ float2 stage[2][kLossStage];
  const bool table = n_fft <= kLossTable, staged = F <= kLossStage;
  const float2 *a = sim + (size_t)blockIdx.x * F;
  const float2 *b = tgt + (size_t)blockIdx.x * F;
  if (table)
    for (int m = threadIdx.x; m < n_fft; m += kBlock) {
      double sn, cs;
      sincospi(2.0 * m / n_fft, &sn, &cs);
      tw[m] = make_double2(cs, sn);
    }
  if (staged)
    for (int n = threadIdx.x; n < F; n += kBlock) {
      stage[0][n] = a[n];
      stage[1][n] = b[n];
    }
  __syncthreads();
  if (staged) {
    a = stage[0];
    b = stage[1];
  }
  const int nb = hi - lo;
  for (int k = lo + threadIdx.x; k < hi; k += kBlock) {
    double ar = 0, ai = 0, br = 0, bi = 0;
    int m = 0;  // k n mod n_fft
    for (int n = 0; n < F; ++n) {
      double s, cth;
      if (table) {
        const double2 t = tw[m];
        cth = t.x;
        s = t.y;
      } else {
        sincospi(2.0 * m / n_fft, &s, &cth);  // exp(-j 2 pi k n / N) = cth - j s
      }
      const float2 x = a[n], y = b[n];
      ar += x.x * cth + x.y * s;
      ai += x.y * cth - x.x * s;
      br += y.x * cth + y.y * s;
      bi += y.y * cth - y.x * s;
      m += k;
      if (m >= n_fft) m -= n_fft;
    }
    mag[0][k - lo] = (float)sqrt(ar * ar + ai * ai);
    mag[1][k - lo] = k < hi_target ? (float)sqrt(br * br + bi * bi) : 0.0f;
  }
  __syncthreads();
  double na = 0, nbv = 0;
  for (int k = threadIdx.x; k < nb; k += kBlock) {
    na += (double)mag[0][k] * mag[0][k];
    nbv += (double)mag[1][k] * mag[1][k];
  }
  sh[0][threadIdx.x] = na;
  sh[1][threadIdx.x] = nbv;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  const double ia = 1.0 / sqrt(sh[0][0]), ib = 1.0 / sqrt(sh[1][0]);
  __syncthreads();
  double l = 0;
  for (int k = threadIdx.x; k < nb; k += kBlock) l += fabs(mag[0][k] * ia - mag[1][k] * ib);
  sh[2][threadIdx.x] = l;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = (float)sh[2][0];
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Classes of the prefix's live lists: about four rays each, at most 65536.
int live_classes(int n) {
  const int64_t c = ((int64_t)n + 3) / 4;
  return (int)std::max<int64_t>(1, std::min<int64_t>(c, 65536));
}

// The synthesis stage's buffers for n_sim simulations of up to m round trips
// each: the dense hit lists, the compaction's tile offsets (the hit count of
// simulation s at off[s][nb] and hits[s]) and the chunks' partials.
struct SynthLayout {
  double *dense;
  int *off, *hits;
  float2 *part;
  size_t bytes;
};

int synth_tiles(int64_t m) { return (int)((m + kCompactTile - 1) / kCompactTile); }

SynthLayout synth_layout(int64_t m, int samples, int n_sim, char *base) {
  SynthLayout L;
  size_t o = 0;
  auto take = [&](size_t b) {
    char *q = base ? base + o : nullptr;
    o += align256(b);
    return q;
  };
  L.dense = (double *)take(sizeof(double) * (size_t)m * n_sim);
  L.off = (int *)take(sizeof(int) * (size_t)(synth_tiles(m) + 1) * n_sim);
  L.hits = (int *)take(sizeof(int) * (size_t)n_sim);
  L.part = (float2 *)take(sizeof(float2) * (size_t)kMaxChunks * samples * n_sim);
  L.bytes = o;
  return L;
}

// Channels of n_sim simulations whose round trips are src[s * stride + i]
// for i below the simulation's length (len[s * len_stride] capped at stride,
// or stride when len is NULL): compaction, synthesis, reduction.
void launch_synth(hipStream_t st, float2 *channels, const double *src, int64_t stride,
                  const int *len, int64_t len_stride, int n_sim, double f0, double df,
                  int samples, const SynthLayout &L) {
  const int nb = synth_tiles(stride);
  hipLaunchKernelGGL(k_hit_count, dim3((unsigned)nb, (unsigned)n_sim), dim3(kBlock), 0, st,
                     L.off, src, stride, len, len_stride, nb);
  hipLaunchKernelGGL(k_live_scan, dim3((unsigned)n_sim), dim3(kScanThreads), 0, st, L.off,
                     L.hits, nb);
  hipLaunchKernelGGL(k_hit_fill, dim3((unsigned)nb, (unsigned)n_sim), dim3(kBlock), 0, st,
                     L.dense, (const int *)L.off, src, stride, len, len_stride, nb);
  const int nfb = (samples + kK - 1) / kK;
  const int items = kMaxChunks * nfb;
  const double inv_c = 1.0 / MMN_C0;
  hipLaunchKernelGGL(k_synth_dense,
                     dim3((unsigned)((items + kWavesPerBlock - 1) / kWavesPerBlock),
                          (unsigned)n_sim),
                     dim3(kBlock), 0, st, L.part, (const double *)L.dense, stride,
                     (const int *)L.hits, (int64_t)1, f0 * inv_c, df * inv_c, samples, nfb);
  hipLaunchKernelGGL(k_synth_reduce, dim3(mmn_blocks(samples, kReduceBlock), (unsigned)n_sim),
                     dim3(kReduceBlock), 0, st, channels, (const float2 *)L.part,
                     (const int *)L.hits, (int64_t)1, samples);
}

bool geom_from(RayGeom &g, int width, int height, float fov_w, float fov_h,
               const float *bbox, const int *res, const float *eye, const float *rx_dir,
               float tau) {
  if (width <= 0 || height <= 0 || (int64_t)width * height > INT32_MAX) return false;
  g.fov_w = fov_w;
  g.fov_h = fov_h;
  for (int a = 0; a < 3; ++a) {
    g.bmin[a] = bbox[a];
    g.bmax[a] = bbox[3 + a];
    g.res[a] = res[a];
    g.eye[a] = eye[a];
    g.rxd[a] = rx_dir[a];
    if (res[a] < 2) return false;
  }
  g.width = width;
  g.height = height;
  g.tau = tau;
  const int64_t v = (int64_t)res[0] * res[1] * res[2];
  g.small = v <= ((int64_t)1 << 24);
  // MMN_RAY_FP32_INDEX=1 keeps the FP32 index arithmetic for every grid (the
  // two forms give the same indices; read at each call)
  const char *e = getenv("MMN_RAY_FP32_INDEX");
  if (e && e[0] == '1') g.small = 0;
  return true;
}

}  // namespace

extern "C" int mmn_ray_matching(float *intersection_pos, float *voxel_position,
                                float *directions, const float *grid, int width,
                                int height, float width_fov, float height_fov,
                                const float *bounding_box, const int *grid_res,
                                const float *eye, const float *rx_dir, void *stream) {
  MMN_REQUIRE(intersection_pos && voxel_position && directions && grid && bounding_box &&
                  grid_res && eye && rx_dir,
              MMN_ERR_NULL, "mmn_ray_matching: NULL pointer argument");
  RayGeom g;
  MMN_REQUIRE(geom_from(g, width, height, width_fov, height_fov, bounding_box, grid_res, eye,
                        rx_dir, 0.0f),
              MMN_ERR_ARG, "mmn_ray_matching: bad sizes");
  const int n = width * height;
  hipLaunchKernelGGL(k_ray_match, dim3(mmn_blocks(n, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, intersection_pos, voxel_position, directions, grid,
                     g);
  return mmn_check_launch("mmn_ray_matching");
}

extern "C" size_t mmn_simulate_channel_workspace_size(int width, int height, int samples) {
  if (width <= 0 || height <= 0 || samples <= 0 || (int64_t)width * height > INT32_MAX)
    return 0;
  const int n = width * height;
  return align256(sizeof(double) * (size_t)n) + synth_layout(n, samples, 1, nullptr).bytes;
}

extern "C" int mmn_simulate_channel(float *channel, const float *grid, float tau, int width,
                                    int height, float width_fov, float height_fov,
                                    const float *bounding_box, const int *grid_res,
                                    const float *eye, const float *rx_dir,
                                    const double *rx_loc, float voxel_size, double f0,
                                    double df, int samples, int is_ti_radar,
                                    void *workspace, size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(channel && grid && bounding_box && grid_res && eye && rx_dir && rx_loc,
              MMN_ERR_NULL, "mmn_simulate_channel: NULL pointer argument");
  RayGeom g;
  MMN_REQUIRE(geom_from(g, width, height, width_fov, height_fov, bounding_box, grid_res, eye,
                        rx_dir, tau) &&
                  samples > 0,
              MMN_ERR_ARG, "mmn_simulate_channel: bad sizes");
  const size_t need = mmn_simulate_channel_workspace_size(width, height, samples);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_simulate_channel: workspace of %zu bytes needed", need);
  const int n = width * height;
  double *dist = (double *)workspace;
  const SynthLayout L =
      synth_layout(n, samples, 1, (char *)workspace + align256(sizeof(double) * (size_t)n));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ray_dist<false>, dim3(mmn_blocks(n, kBlock)), dim3(kBlock), 0, st, dist,
                     grid, g, voxel_size, rx_loc[0], rx_loc[1], rx_loc[2], is_ti_radar,
                     (const SimParam *)nullptr, (int64_t)0);
  launch_synth(st, (float2 *)channel, dist, n, nullptr, 0, 1, f0, df, samples, L);
  return mmn_check_launch("mmn_simulate_channel");
}

namespace {

size_t batch_ws(int n, int samples, int nsim) {
  return align256(sizeof(SimParam) * (size_t)nsim) +
         align256(sizeof(double) * (size_t)n * nsim) +
         synth_layout(n, samples, nsim, nullptr).bytes;
}

}  // namespace

extern "C" size_t mmn_simulate_channels_workspace_size(int width, int height, int samples,
                                                       int n_sim) {
  if (width <= 0 || height <= 0 || samples <= 0 || n_sim <= 0 ||
      (int64_t)width * height > INT32_MAX)
    return 0;
  return batch_ws(width * height, samples, n_sim);
}

extern "C" int mmn_simulate_channels(float *channels, const float *grid, int n_comp,
                                     const int32_t *comp, const float *taus,
                                     const float *eyes, const float *rx_dirs,
                                     const double *rx_locs, int n_sim, int width, int height,
                                     float width_fov, float height_fov,
                                     const float *bounding_box, const int *grid_res,
                                     float voxel_size, double f0, double df, int samples,
                                     int is_ti_radar, void *workspace, size_t workspace_bytes,
                                     void *stream) {
  MMN_REQUIRE(n_sim >= 0 && n_sim <= kMaxSims, MMN_ERR_ARG,
              "mmn_simulate_channels: n_sim must be in [0, %d] (got %d)", kMaxSims, n_sim);
  if (n_sim == 0) return 0;
  MMN_REQUIRE(channels && grid && comp && taus && eyes && rx_dirs && rx_locs && bounding_box &&
                  grid_res,
              MMN_ERR_NULL, "mmn_simulate_channels: NULL pointer argument");
  const float zero3[3] = {0.0f, 0.0f, 0.0f};
  RayGeom g;
  MMN_REQUIRE(geom_from(g, width, height, width_fov, height_fov, bounding_box, grid_res, zero3,
                        zero3, 0.0f) &&
                  samples > 0 && n_comp > 0,
              MMN_ERR_ARG, "mmn_simulate_channels: bad sizes");
  for (int s = 0; s < n_sim; ++s)
    MMN_REQUIRE(comp[s] >= 0 && comp[s] < n_comp, MMN_ERR_ARG,
                "mmn_simulate_channels: comp[%d] = %d outside [0, %d)", s, comp[s], n_comp);
  const int n = width * height;
  const size_t need = batch_ws(n, samples, n_sim);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_simulate_channels: workspace of %zu bytes needed", need);
  // the simulations' parameters, packed on the host and copied to the
  // workspace's head (the copy has left the host buffer when it returns)
  std::vector<SimParam> prm((size_t)n_sim);
  for (int s = 0; s < n_sim; ++s) {
    SimParam &q = prm[(size_t)s];
    for (int a = 0; a < 3; ++a) {
      q.eye[a] = eyes[3 * s + a];
      q.rxd[a] = rx_dirs[3 * s + a];
      q.rx[a] = rx_locs[3 * s + a];
    }
    q.tau = taus[s];
    q.comp = comp[s];
    q.pad = 0.0;
  }
  char *w = (char *)workspace;
  SimParam *sp = (SimParam *)w;
  double *dist = (double *)(w + align256(sizeof(SimParam) * (size_t)n_sim));
  const SynthLayout L = synth_layout(n, samples, n_sim,
                                     (char *)dist + align256(sizeof(double) * (size_t)n * n_sim));
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(sp, prm.data(), sizeof(SimParam) * (size_t)n_sim,
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    mmn_set_error("mmn_simulate_channels: parameter copy: %s", hipGetErrorString(e));
    return (int)e;
  }
  const int64_t vox = (int64_t)grid_res[0] * grid_res[1] * grid_res[2];
  hipLaunchKernelGGL(k_ray_dist<true>, dim3(mmn_blocks(n, kBlock), (unsigned)n_sim),
                     dim3(kBlock), 0, st, dist, grid, g, voxel_size, 0.0, 0.0, 0.0, is_ti_radar,
                     (const SimParam *)sp, vox);
  launch_synth(st, (float2 *)channels, dist, n, nullptr, 0, n_sim, f0, df, samples, L);
  return mmn_check_launch("mmn_simulate_channels");
}

namespace {

struct PrefixLayout {
  SimParam *sp;
  RayState *st;
  int *off, *list, *tot;
  size_t bytes;
};

PrefixLayout prefix_layout(int n, int n_sim, char *base) {
  const int64_t Q = live_classes(n);
  PrefixLayout L;
  size_t o = 0;
  auto take = [&](size_t b) {
    char *q = base ? base + o : nullptr;
    o += align256(b);
    return q;
  };
  L.sp = (SimParam *)take(sizeof(SimParam) * (size_t)n_sim);
  L.st = (RayState *)take(sizeof(RayState) * (size_t)n * n_sim);
  L.off = (int *)take(sizeof(int) * (size_t)(Q + 1) * n_sim);
  L.list = (int *)take(sizeof(int) * (size_t)n * n_sim);
  L.tot = (int *)take(sizeof(int) * (size_t)n_sim);
  L.bytes = o;
  return L;
}

// Host checks shared by the batched entry points.
int check_batch(const char *what, int n_sim, int n_comp, const int32_t *comp,
                const float *bounding_box, const int *grid_res, int width, int height,
                float width_fov, float height_fov, RayGeom &g) {
  MMN_REQUIRE(n_sim > 0 && n_sim <= kMaxSims, MMN_ERR_ARG, "%s: n_sim must be in [1, %d]",
              what, kMaxSims);
  const float zero3[3] = {0.0f, 0.0f, 0.0f};
  MMN_REQUIRE(n_comp > 0 && geom_from(g, width, height, width_fov, height_fov, bounding_box,
                                      grid_res, zero3, zero3, 0.0f),
              MMN_ERR_ARG, "%s: bad sizes", what);
  for (int s = 0; s < n_sim; ++s)
    MMN_REQUIRE(comp[s] >= 0 && comp[s] < n_comp, MMN_ERR_ARG,
                "%s: comp[%d] = %d outside [0, %d)", what, s, comp[s], n_comp);
  return 0;
}

int copy_params(const char *what, SimParam *dst, int n_sim, const int32_t *comp,
                const float *taus, const float *eyes, const float *rx_dirs,
                const double *rx_locs, hipStream_t st) {
  std::vector<SimParam> prm((size_t)n_sim);
  for (int s = 0; s < n_sim; ++s) {
    SimParam &q = prm[(size_t)s];
    for (int a = 0; a < 3; ++a) {
      q.eye[a] = eyes[3 * s + a];
      q.rxd[a] = rx_dirs[3 * s + a];
      q.rx[a] = rx_locs ? rx_locs[3 * s + a] : 0.0;
    }
    q.tau = taus ? taus[s] : 0.0f;
    q.comp = comp[s];
    q.pad = 0.0;
  }
  const hipError_t e = hipMemcpyAsync(dst, prm.data(), sizeof(SimParam) * (size_t)n_sim,
                                      hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    mmn_set_error("%s: parameter copy: %s", what, hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // namespace

extern "C" size_t mmn_ray_prefix_size(int width, int height, int n_sim) {
  if (width <= 0 || height <= 0 || n_sim <= 0 || (int64_t)width * height > INT32_MAX)
    return 0;
  return prefix_layout(width * height, n_sim, nullptr).bytes;
}

extern "C" int mmn_ray_prefix(void *prefix, size_t prefix_bytes, int *live_counts,
                              const float *grid, int n_comp, const int32_t *comp,
                              const float *eyes, const float *rx_dirs, int n_sim, int width,
                              int height, float width_fov, float height_fov,
                              const float *bounding_box, const int *grid_res, float tau_lo,
                              void *stream) {
  MMN_REQUIRE(prefix && live_counts && grid && comp && eyes && rx_dirs && bounding_box &&
                  grid_res,
              MMN_ERR_NULL, "mmn_ray_prefix: NULL pointer argument");
  RayGeom g;
  int rc = check_batch("mmn_ray_prefix", n_sim, n_comp, comp, bounding_box, grid_res, width,
                       height, width_fov, height_fov, g);
  if (rc) return rc;
  MMN_REQUIRE(std::isfinite(tau_lo), MMN_ERR_ARG, "mmn_ray_prefix: tau_lo must be finite");
  const int n = width * height;
  PrefixLayout L = prefix_layout(n, n_sim, (char *)prefix);
  MMN_REQUIRE(prefix_bytes >= L.bytes, MMN_ERR_WS,
              "mmn_ray_prefix: prefix buffer of %zu bytes needed", L.bytes);
  hipStream_t st = (hipStream_t)stream;
  rc = copy_params("mmn_ray_prefix", L.sp, n_sim, comp, nullptr, eyes, rx_dirs, nullptr, st);
  if (rc) return rc;
  const int64_t vox = (int64_t)grid_res[0] * grid_res[1] * grid_res[2];
  const int Q = live_classes(n);
  hipLaunchKernelGGL(k_ray_prefix, dim3(mmn_blocks(n, kBlock), (unsigned)n_sim), dim3(kBlock),
                     0, st, L.st, grid, g, (const SimParam *)L.sp, vox, tau_lo);
  hipLaunchKernelGGL(k_live_count, dim3(mmn_blocks(Q, kBlock), (unsigned)n_sim), dim3(kBlock),
                     0, st, L.off, (const RayState *)L.st, n, Q);
  hipLaunchKernelGGL(k_live_scan, dim3((unsigned)n_sim), dim3(kScanThreads), 0, st, L.off,
                     L.tot, Q);
  hipLaunchKernelGGL(k_live_fill, dim3(mmn_blocks(Q, kBlock), (unsigned)n_sim), dim3(kBlock),
                     0, st, L.list, (const int *)L.off, (const RayState *)L.st, n, Q);
  rc = mmn_check_launch("mmn_ray_prefix");
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(live_counts, L.tot, sizeof(int) * (size_t)n_sim,
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    mmn_set_error("mmn_ray_prefix: live counts: %s", hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

extern "C" size_t mmn_simulate_channels_prefixed_workspace_size(int width, int height,
                                                                int samples, int n_sim,
                                                                int live_max) {
  if (width <= 0 || height <= 0 || samples <= 0 || n_sim <= 0 || live_max < 0 ||
      (int64_t)width * height > INT32_MAX)
    return 0;
  const int lm = std::max(live_max, 1);
  return align256(sizeof(SimParam) * (size_t)n_sim) +
         align256(sizeof(double) * (size_t)lm * n_sim) +
         synth_layout(lm, samples, n_sim, nullptr).bytes;
}

extern "C" int mmn_simulate_channels_prefixed(
    float *channels, const void *prefix, size_t prefix_bytes, const int *live_counts,
    const float *grid, int n_comp, const int32_t *comp, const float *taus, const float *eyes,
    const float *rx_dirs, const double *rx_locs, int n_sim, int width, int height,
    float width_fov, float height_fov, const float *bounding_box, const int *grid_res,
    float voxel_size, double f0, double df, int samples, int is_ti_radar, float tau_lo,
    void *workspace, size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(channels && prefix && live_counts && grid && comp && taus && eyes && rx_dirs &&
                  rx_locs && bounding_box && grid_res,
              MMN_ERR_NULL, "mmn_simulate_channels_prefixed: NULL pointer argument");
  RayGeom g;
  int rc = check_batch("mmn_simulate_channels_prefixed", n_sim, n_comp, comp, bounding_box,
                       grid_res, width, height, width_fov, height_fov, g);
  if (rc) return rc;
  MMN_REQUIRE(samples > 0, MMN_ERR_ARG, "mmn_simulate_channels_prefixed: samples <= 0");
  const int n = width * height;
  int live_max = 0;
  for (int s = 0; s < n_sim; ++s) {
    MMN_REQUIRE(taus[s] >= tau_lo, MMN_ERR_ARG,
                "mmn_simulate_channels_prefixed: taus[%d] below the prefix's tau_lo", s);
    MMN_REQUIRE(live_counts[s] >= 0 && live_counts[s] <= n, MMN_ERR_ARG,
                "mmn_simulate_channels_prefixed: live_counts[%d] outside [0, rays]", s);
    live_max = std::max(live_max, live_counts[s]);
  }
  const PrefixLayout L = prefix_layout(n, n_sim, (char *)prefix);
  MMN_REQUIRE(prefix_bytes >= L.bytes, MMN_ERR_WS,
              "mmn_simulate_channels_prefixed: prefix of %zu bytes expected", L.bytes);
  const size_t need = mmn_simulate_channels_prefixed_workspace_size(width, height, samples,
                                                                    n_sim, live_max);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_simulate_channels_prefixed: workspace of %zu bytes needed", need);
  char *w = (char *)workspace;
  SimParam *sp = (SimParam *)w;
  const int lm = std::max(live_max, 1);
  double *dl = (double *)(w + align256(sizeof(SimParam) * (size_t)n_sim));
  const SynthLayout S =
      synth_layout(lm, samples, n_sim, (char *)dl + align256(sizeof(double) * (size_t)lm * n_sim));
  hipStream_t st = (hipStream_t)stream;
  rc = copy_params("mmn_simulate_channels_prefixed", sp, n_sim, comp, taus, eyes, rx_dirs,
                   rx_locs, st);
  if (rc) return rc;
  const int64_t vox = (int64_t)grid_res[0] * grid_res[1] * grid_res[2];
  const int Q = live_classes(n);
  hipLaunchKernelGGL(k_ray_resume, dim3(mmn_blocks(lm, kBlock), (unsigned)n_sim), dim3(kBlock),
                     0, st, dl, grid, g, voxel_size, is_ti_radar, (const SimParam *)sp, vox,
                     (const RayState *)L.st, (const int *)L.list, (const int *)L.off, Q, lm);
  // the live rays' round trips (NaN: no hit at this offset), in ray order;
  // simulation s has off[s][Q] of them
  launch_synth(st, (float2 *)channels, dl, lm, (const int *)L.off + Q, (int64_t)Q + 1, n_sim,
               f0, df, samples, S);
  return mmn_check_launch("mmn_simulate_channels_prefixed");
}

extern "C" int mmn_channel_loss(float *loss, const float *sim, const float *target,
                                int n_pairs, int samples, int n_fft, int bin_lo, int bin_hi,
                                int bin_hi_target, void *stream) {
  MMN_REQUIRE(loss && sim && target, MMN_ERR_NULL, "mmn_channel_loss: NULL pointer argument");
  MMN_REQUIRE(n_pairs >= 0 && samples > 0 && n_fft >= samples && bin_lo >= 0 &&
                  bin_lo < bin_hi && bin_hi <= n_fft && bin_hi - bin_lo <= 1024,
              MMN_ERR_ARG, "mmn_channel_loss: bad sizes or bins");
  if (n_pairs == 0) return 0;
  hipLaunchKernelGGL(k_channel_loss, dim3(n_pairs), dim3(kBlock), 0, (hipStream_t)stream, loss,
                     (const float2 *)sim, (const float2 *)target, samples, n_fft, bin_lo,
                     bin_hi, bin_hi_target);
  return mmn_check_launch("mmn_channel_loss");
}
