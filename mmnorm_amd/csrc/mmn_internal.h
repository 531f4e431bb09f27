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

// Zeroes bytes (whole 32-bit words, 4-byte aligned) at p on stream st with a
// fill kernel, not hipMemsetAsync, so that callers may capture the entry
// points in a HIP graph; returns 0 or an error code (message under what).
int mmn_zero_async(void *p, size_t bytes, hipStream_t st, const char *what);

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

// exp(-j 2 pi t) with the phase given in quarter turns, t4 = 4 t.  The
// integer quarter count q = rint(t4) and the remainder y = t4 - q (exact in
// FP64, |y| <= 1/2) are split off in FP64; the rest is FP32: theta = y pi/2
// in [-pi/4, pi/4], sin and cos by their Taylor series through theta^9 and
// theta^10 (truncation below 2.5e-10 there), and the quadrant applied as a
// swap and two sign flips:
//   2 pi t = q pi/2 + theta,  q mod 4 = 0: ( c,  s)  1: (-s,  c)
//                                       2: (-c, -s)  3: ( s, -c)  as (cos, sin).
// q is converted to int32 directly (one instruction): |t4| < 2^31, which
// holds for every step phase here (t4 = 4 d df/c is below 10^3 for d < 1 km
// at df = 75 MHz).
__device__ __forceinline__ void phasor_quarter(double t4, float &re, float &im) {
  const double qd = rint(t4);
  const float y = (float)(t4 - qd);
  const int q = (int)qd;
  const float th = y * 1.57079632679489662f;
  const float t2 = th * th;
  float ps = fmaf(t2, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(t2, ps, 8.33333333e-3f);
  ps = fmaf(t2, ps, -1.66666667e-1f);
  const float sn = fmaf(th * t2, ps, th);
  float pc = fmaf(t2, -2.75573192e-7f, 2.48015873e-5f);
  pc = fmaf(t2, pc, -1.38888889e-3f);
  pc = fmaf(t2, pc, 4.16666667e-2f);
  pc = fmaf(t2, pc, -0.5f);
  const float cs = fmaf(t2, pc, 1.0f);
  const bool swap = q & 1;
  const float cv = swap ? sn : cs;
  const float sv = swap ? cs : sn;
  // cos negative for q = 1, 2; sin negative for q = 2, 3; im = -sin.
  const unsigned cneg = (unsigned)((q + 1) & 2) << 30;
  const unsigned sneg = ((unsigned)(q & 2) << 30) ^ 0x80000000u;
  re = __uint_as_float(__float_as_uint(cv) ^ cneg);
  im = __uint_as_float(__float_as_uint(sv) ^ sneg);
}

// exp(-j 2 pi t) for t in revolutions: the integer part is removed in FP64
// and the remaining fraction (|fr| <= 1/2) goes to the hardware sine and
// cosine, which take their argument in revolutions (measured on MI355X over
// [-1/2, 1/2]: max abs error 1.3e-7, rms 3.5e-8).
__device__ __forceinline__ void phasor_hw(double t, float &re, float &im) {
  const float fr = (float)(t - rint(t));
  re = __builtin_amdgcn_cosf(fr);
  im = -__builtin_amdgcn_sinf(fr);
}

// exp(-j 2 pi u) for a small FP32 phase u (revolutions), no FP64.
__device__ __forceinline__ void phasor_rev32(float u, float &re, float &im) {
  float s, c;
  sincospif(2.0f * u, &s, &c);
  re = c;
  im = -s;
}

// Adds one row's contribution c to the image (Kahan) and the vote moment.
__device__ __forceinline__ void add_row(float cr, float ci, float vecx,
                                        float vecy, float vecz, float &Ir,
                                        float &Ii, float &Cr, float &Ci,
                                        float (&M)[6]) {
  // Kahan: C carries the low-order part lost by the running sum.
  const float yr = cr - Cr, yi = ci - Ci;
  const float tr = Ir + yr, ti = Ii + yi;
  Cr = (tr - Ir) - yr;
  Ci = (ti - Ii) - yi;
  Ir = tr;
  Ii = ti;
  M[0] = fmaf(vecx, cr, M[0]);
  M[1] = fmaf(vecx, ci, M[1]);
  M[2] = fmaf(vecy, cr, M[2]);
  M[3] = fmaf(vecy, ci, M[3]);
  M[4] = fmaf(vecz, cr, M[4]);
  M[5] = fmaf(vecz, ci, M[5]);
}

// Fixed-order FP64 sum of rs row segments' partial images (pimg, rs x n) and
// moments (pmom, rs x n x 3, may be null) into img / mom (bp_forward.hip).
void mmn_rs_reduce(hipStream_t st, float2 *img, float2 *mom, const float2 *pimg,
                   const float2 *pmom, int64_t n, int rs);

static inline unsigned mmn_blocks(int64_t n, int per_block) {
  return (unsigned)((n + per_block - 1) / per_block);
}
