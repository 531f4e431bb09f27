// Internal helpers shared by the HIP translation units of libmmnorm_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "mmnorm_amd.h"

// Speed of light used by the reference's phase (imaging_gpu.cu:3).
#define MMN_C0 2.99792458e8

// ---------------------------------------------------------------- errors --
// Thread-local last-error message (defined in capi.hip).
void mmn_set_error(const char *fmt, ...);

#define MMN_REQUIRE(cond, code, ...)        \
  do {                                      \
    if (!(cond)) {                          \
      mmn_set_error(__VA_ARGS__);           \
      return (code);                        \
    }                                       \
  } while (0)

// Checks the launch that was just enqueued; returns the hipError_t (> 0).
int mmn_check_launch(const char *what);

// ------------------------------------------------------------- geometry --
// Linear voxel index -> (x, y, z) with z fastest (imaging_gpu.cu:48-54).
struct VoxelGrid {
  const float *__restrict__ xl;
  const float *__restrict__ yl;
  const float *__restrict__ zl;
  int ny, nz;
};

__device__ __forceinline__ void voxel_coords(const float *__restrict__ xl,
                                             const float *__restrict__ yl,
                                             const float *__restrict__ zl,
                                             int ny, int nz, int64_t lin,
                                             float &vx, float &vy, float &vz) {
  int64_t yz = (int64_t)ny * nz;
  int x = (int)(lin / yz);
  int64_t rem = lin - (int64_t)x * yz;
  int y = (int)(rem / nz);
  int z = (int)(rem - (int64_t)y * nz);
  vx = xl[x];
  vy = yl[y];
  vz = zl[z];
}

// Round-trip distance |v - p| + |v - p - o| evaluated in FP64 from FP32
// inputs.  The squared norms are formed in FP64 (the differences of two
// floats are exact there); sqrt uses the hardware FP32 reciprocal square
// root as a seed and one FP64 Newton step, which leaves a relative error
// near 1e-14 at a fraction of the cost of a correctly rounded FP64 sqrt.
__device__ __forceinline__ double sqrt_nr64(double s) {
  double y = (double)__builtin_amdgcn_rsqf((float)s);  // ~1 ulp of FP32
  double d0 = s * y;
  double e = fma(-d0, d0, s);
  return fma(0.5 * y, e, d0);
}

__device__ __forceinline__ double roundtrip_dist64(float vx, float vy,
                                                   float vz, float px,
                                                   float py, float pz,
                                                   float ox, float oy,
                                                   float oz, bool mono) {
  double dx = (double)vx - (double)px;
  double dy = (double)vy - (double)py;
  double dz = (double)vz - (double)pz;
  double s1 = fma(dz, dz, fma(dy, dy, dx * dx));
  double d1 = sqrt_nr64(s1);
  if (mono) return d1 + d1;
  double ex = dx - (double)ox, ey = dy - (double)oy, ez = dz - (double)oz;
  double s2 = fma(ez, ez, fma(ey, ey, ex * ex));
  return d1 + sqrt_nr64(s2);
}

// exp(-j 2 pi t) for a phase t given in revolutions, with the integer part
// removed in FP64 (exact for the |t| < 2^52 seen here) before the FP32
// evaluation: the only FP64 on the recurrence path.
__device__ __forceinline__ void phasor_rev64(double t, float &re, float &im) {
  double fr = t - rint(t);  // in [-0.5, 0.5]
  float s, c;
  sincospif(2.0f * (float)fr, &s, &c);
  re = c;
  im = -s;
}

// exp(-j 2 pi u) for a small FP32 phase u (revolutions), no FP64.
__device__ __forceinline__ void phasor_rev32(float u, float &re, float &im) {
  float s, c;
  sincospif(2.0f * u, &s, &c);
  re = c;
  im = -s;
}

static inline unsigned mmn_blocks(int64_t n, int per_block) {
  return (unsigned)((n + per_block - 1) / per_block);
}
