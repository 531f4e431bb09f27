// Backprojection kernels for gfx950 (MI355X, wave64).
//
//   k_image_mirror  the reference kernel's arithmetic (imaging_gpu.cu:48-122):
//                   FP32 distance, FP64 phase and cos/sin, FP32 sums in row
//                   order; both the interpolated and the full branch.
//   k_bp_plain      full matched filter with one FP32 sine/cosine pair per
//                   (voxel, row, sample): the direct baselines, with the
//                   library sincosf of the phase in radians, or (FAST) the
//                   hardware sine/cosine of its fraction in revolutions.
//   k_bp_rec        full matched filter evaluated per (voxel, row) as a
//                   polynomial in the per-sample phase step r = exp(-j 2pi df d/c)
//                   (Horner: one complex rotate-and-add, 4 FMAs, per sample),
//                   re-anchored every K samples by an FP64-reduced phasor,
//                   rows summed with Kahan compensation, and the normal-vote
//                   moment accumulated in the same pass.
//
// Thread mapping (all three): one lane per voxel, 256-lane workgroups; every
// lane of a wave walks the same row at the same time, so the row's antenna
// position, RX offset and measurement samples are wave-uniform and are read
// through scalar loads straight into SGPR operands of the FMAs -- no LDS
// staging and no per-lane measurement traffic.
#include <stdlib.h>

#include <algorithm>

#include "mmn_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kBlock = 256;

// --------------------------------------------------------------------------
// Reference-arithmetic mirror.  Contraction is off so every product and sum
// rounds separately, as the restatement in oracle/ does; the operation order
// of each expression follows imaging_gpu.cu:77-117.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_image_mirror(
    float2 *__restrict__ img, const float *__restrict__ xl,
    const float *__restrict__ yl, const float *__restrict__ zl,
    const float *__restrict__ ant, const float2 *__restrict__ meas,
    const float *__restrict__ rxo, float slope, float wavelength,
    float fft_spacing, int ny, int nz, int64_t nvox, int rows, int R, int S,
    int is_ti, int interp) {
#pragma clang fp contract(off)
  const int64_t lin = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (lin >= nvox) return;
  float vx, vy, vz;
  voxel_coords(xl, yl, zl, ny, nz, lin, vx, vy, vz);

  float sum_r = 0.0f, sum_i = 0.0f;
  for (int i = 0; i < rows; ++i) {
    const float ax = vx - ant[3 * i + 0];
    const float ay = vy - ant[3 * i + 1];
    const float az = vz - ant[3 * i + 2];
    const int rx = i % R;
    const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
    const float fwd = sqrtf(ax * ax + ay * ay + az * az);
    const float bx = ax - ox, by = ay - oy, bz = az - oz;
    const float back = sqrtf(bx * bx + by * by + bz * bz);
    float d = fwd + back;
    if (is_ti) d = (float)((double)d + 0.15);
    const float2 *__restrict__ row = meas + (size_t)i * S;
    if (interp) {
      if (d < 0.0f || d > fft_spacing * (float)S) continue;
      const int bin = (int)floorf(d / fft_spacing / 2.0f);
      const float2 m = row[bin];
      const double th = -2.0 * kPi * (double)d / (double)wavelength;
      const float pr = (float)cos(th);
      const float pi = (float)sin(th);
      sum_r += m.x * pr - m.y * pi;
      sum_i += m.x * pi + m.y * pr;
    } else {
      const double a0 = -2.0 * kPi * (double)d / (double)wavelength;
      for (int j = 0; j < S; ++j) {
        const float2 m = row[j];
        const double th =
            a0 + -2.0 * kPi * (double)d * (double)j * (double)slope / MMN_C0;
        const float pr = (float)cos(th);
        const float pi = (float)sin(th);
        sum_r += m.x * pr - m.y * pi;
        sum_i += m.x * pi + m.y * pr;
      }
    }
  }
  img[lin] = make_float2(sum_r, sum_i);
}

// --------------------------------------------------------------------------
// Shared per-lane voxel set-up for the fast kernels.
// --------------------------------------------------------------------------
template <int VPT>
__device__ __forceinline__ void load_voxels(
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    int64_t nvox, int ny, int nz, int64_t base, float (&vx)[VPT],
    float (&vy)[VPT], float (&vz)[VPT]) {
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    int64_t t = base + (int64_t)k * kBlock;
    if (t >= nvox) t = nvox - 1;  // idle lanes recompute the last voxel
    const int64_t lin = vidx ? (int64_t)vidx[t] : t;
    voxel_coords(xl, yl, zl, ny, nz, lin, vx[k], vy[k], vz[k]);
  }
}

// Marks the voxel coordinates as ready before a row loop that gathers: the
// wait for their loads then comes once, here.  Otherwise the first use inside
// the loop carries a wait on the load counter, which in every later row also
// waits for the gathers issued before it in that row, so that a lane's
// gathers return one after another instead of together.
template <int VPT>
__device__ __forceinline__ void settle_voxels(float (&vx)[VPT], float (&vy)[VPT],
                                              float (&vz)[VPT]) {
#pragma unroll
  for (int k = 0; k < VPT; ++k) asm volatile("" : "+v"(vx[k]), "+v"(vy[k]), "+v"(vz[k]));
}

// Voxel t (< nvox) of the launch goes to img[t - off] (and mom[t - off]).
template <int VPT>
__device__ __forceinline__ void store_voxels(
    float2 *__restrict__ img, float2 *__restrict__ mom, int64_t nvox,
    int64_t base, const float (&Ir)[VPT], const float (&Ii)[VPT],
    const float (&M)[VPT][6], int64_t off = 0) {
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    const int64_t t = base + (int64_t)k * kBlock;
    if (t < nvox) {
      const int64_t o = t - off;
      img[o] = make_float2(Ir[k], Ii[k]);
      if (mom) {
        mom[3 * o + 0] = make_float2(M[k][0], M[k][1]);
        mom[3 * o + 1] = make_float2(M[k][2], M[k][3]);
        mom[3 * o + 2] = make_float2(M[k][4], M[k][5]);
      }
    }
  }
}

// --------------------------------------------------------------------------
// Plain: FP32 distance, FP32 phase theta = d * 2 pi f_j / c, sincosf.
// --------------------------------------------------------------------------
template <int VPT, bool FAST>
__global__ __launch_bounds__(kBlock) void k_bp_plain(
    float2 *__restrict__ img, float2 *__restrict__ mom,
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    int64_t nvox, const float *__restrict__ ant,
    const float2 *__restrict__ meas, const float *__restrict__ rxo, float k0,
    float dk, float u0, float du, int ny, int nz, int rows, int R, int F, int is_ti) {
  const int64_t base = (int64_t)blockIdx.x * (kBlock * VPT) + threadIdx.x;
  float vx[VPT], vy[VPT], vz[VPT];
  load_voxels<VPT>(xl, yl, zl, vidx, nvox, ny, nz, base, vx, vy, vz);
  float Ir[VPT], Ii[VPT], M[VPT][6];
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    Ir[k] = Ii[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) M[k][q] = 0.0f;
  }
  for (int i = 0; i < rows; ++i) {
    const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
    const int rx = i % R;
    const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
    const float2 *__restrict__ row = meas + (size_t)i * F;
    float d[VPT], cr[VPT], ci[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const float ax = vx[k] - px, ay = vy[k] - py, az = vz[k] - pz;
      const float bx = ax - ox, by = ay - oy, bz = az - oz;
      d[k] = sqrtf(ax * ax + ay * ay + az * az) + sqrtf(bx * bx + by * by + bz * bz);
      if (is_ti) d[k] += 0.15f;
      cr[k] = ci[k] = 0.0f;
    }
    // one sample for every voxel of the lane
    auto sample = [&](int j) {
      const float2 m = row[j];
      const float kj = fmaf((float)j, dk, k0);
      const float uj = fmaf((float)j, du, u0);
#pragma unroll
      for (int k = 0; k < VPT; ++k) {
        float s, c;
        if (FAST) {
          // phase in revolutions, FP32; its fraction to the hardware sine
          // and cosine (which take revolutions)
          const float t = d[k] * uj;
          const float fr = t - rintf(t);
          s = __builtin_amdgcn_sinf(fr);
          c = __builtin_amdgcn_cosf(fr);
        } else {
          sincosf(d[k] * kj, &s, &c);
        }
        // m * exp(-j theta) = m * (c - j s)
        cr[k] = fmaf(m.x, c, fmaf(m.y, s, cr[k]));
        ci[k] = fmaf(m.y, c, fmaf(-m.x, s, ci[k]));
      }
    };
    // groups of 8 samples, so that their loads go out together (one 64-byte
    // scalar load) instead of one load and one wait per sample
    int j = 0;
    for (; j + 8 <= F; j += 8) {
#pragma unroll
      for (int q = 0; q < 8; ++q) sample(j + q);
    }
    for (; j < F; ++j) sample(j);
    const float vax = px + ox / 2.0f, vay = py + oy / 2.0f, vaz = pz + oz / 2.0f;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      Ir[k] += cr[k];
      Ii[k] += ci[k];
      const float vecx = vax - vx[k], vecy = vay - vy[k], vecz = vaz - vz[k];
      M[k][0] = fmaf(vecx, cr[k], M[k][0]);
      M[k][1] = fmaf(vecx, ci[k], M[k][1]);
      M[k][2] = fmaf(vecy, cr[k], M[k][2]);
      M[k][3] = fmaf(vecy, ci[k], M[k][3]);
      M[k][4] = fmaf(vecz, cr[k], M[k][4]);
      M[k][5] = fmaf(vecz, ci[k], M[k][5]);
    }
  }
  store_voxels<VPT>(img, mom, nvox, base, Ir, Ii, M);
}

// --------------------------------------------------------------------------
// Interpolated branch, MI355X form (k_bp_interp): per (voxel, row) the
// reference's FP32 round-trip distance and range bin, computed with the same
// operations so that every bin is the reference's (imaging_gpu.cu:77-97);
// the phasor exp(-j 2 pi d / lambda) from the FP64-reduced phase and the
// hardware sine/cosine instead of FP64 cos/sin; and the normal-vote moment
// M = sum_i (p_i + o_i/2 - v) c_i in the same pass, so that the vote of
// surface_normal_est.cu:110-149 is Re(conj(I) M) / |I| without a second pass.
// --------------------------------------------------------------------------
// The correctly rounded FP32 square root (round to nearest even, the value
// sqrtf returns) in fewer instructions than sqrtf: the hardware square root
// s (within 1 ulp) and a test of its neighbours by the sign of the residual
// x - s's, the correction the compiler's own sqrtf applies.  sqrtf adds a
// scaling of arguments below 2^-96, where the hardware result is not within
// 1 ulp, and a select for 0 and infinity; the neighbour test already returns
// 0, -0, infinity and NaN unchanged, so only arguments below 2^-96 (squared
// distances under 1.3e-29 m^2, and negative ones) go to sqrtf here.
// mmn_sqrt_rn_check compares the two over every float.
__device__ __forceinline__ float sqrt_rn(float x) {
  if (x < 0x1p-96f) return sqrtf(x);
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  float r = fmaf(-dn, s, x) <= 0.0f ? dn : s;
  r = fmaf(-up, s, x) > 0.0f ? up : r;
  return r;
}

// Bit patterns [first, first + count) as floats: counts where sqrt_rn and
// sqrtf differ (NaN results compare equal whatever their payload) in res[0]
// and keeps the smallest such pattern in res[1].
__global__ __launch_bounds__(kBlock) void k_sqrt_rn_check(unsigned long long *__restrict__ res,
                                                          uint64_t first, uint64_t count) {
  unsigned long long bad = 0, lo = ~0ull;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * kBlock) {
    const uint32_t bits = (uint32_t)(first + t);
    const float x = __uint_as_float(bits);
    const float a = sqrt_rn(x), b = sqrtf(x);
    const bool same = __float_as_uint(a) == __float_as_uint(b) || (isnan(a) && isnan(b));
    if (!same) {
      ++bad;
      lo = min(lo, (unsigned long long)bits);
    }
  }
  if (bad) {
    atomicAdd(&res[0], bad);
    atomicMin(&res[1], lo);
  }
}

__device__ __forceinline__ float interp_dist(float vx, float vy, float vz, float px, float py,
                                             float pz, float ox, float oy, float oz, bool mono,
                                             int is_ti) {
#pragma clang fp contract(off)
  const float ax = vx - px, ay = vy - py, az = vz - pz;
  const float fwd = sqrt_rn(ax * ax + ay * ay + az * az);
  float back = fwd;  // o = 0: the second leg is the first, bit for bit
  if (!mono) {
    const float bx = ax - ox, by = ay - oy, bz = az - oz;
    back = sqrt_rn(bx * bx + by * by + bz * bz);
  }
  float d = fwd + back;
  if (is_ti) d = (float)((double)d + 0.15);
  return d;
}

__device__ __forceinline__ int interp_bin(float d, float fft_spacing) {
#pragma clang fp contract(off)
  return (int)floorf(d / fft_spacing / 2.0f);
}

// floor(RN(d / fs) / 2) without the correctly rounded division where it is
// not needed: q = d * RN(1 / fs) is within 2 ulp of RN(d / fs), so when the
// interval q * (1 -+ 4e-7) (3.3 ulp each side) holds no even integer the
// floor is decided; otherwise the exact division runs.  The halving is
// folded into the factors: RN(q * RN(1 -+ 4e-7)) / 2 = RN(q * (RN(1 -+ 4e-7) / 2))
// exactly, a power-of-two scale commuting with rounding (no underflow here).
__device__ __forceinline__ int interp_bin_fast(float d, float fft_spacing, float inv_fs) {
#pragma clang fp contract(off)
  const float q = d * inv_fs;
  const float a = floorf(q * ((1.0f - 4e-7f) * 0.5f));
  const float b = floorf(q * ((1.0f + 4e-7f) * 0.5f));
  if (a == b) return (int)a;
  return interp_bin(d, fft_spacing);
}

__device__ __forceinline__ float half_off(float p, float o) {
#pragma clang fp contract(off)
  return p + o / 2.0f;
}

// exp(-j 2 pi d / lambda) for an FP32 distance d, with 1/lambda split into
// two floats hi + lo (hi = RN(1/lambda), lo = RN(1/lambda - hi)): d * hi =
// p + e exactly (e from an FMA), so the phase's fraction
// (p - rint(p)) + (e + d * lo) is formed in FP32 to about an ulp of the
// fraction (p - rint(p) is exact), as the FP64 product and reduction gives
// it, and goes to the hardware sine and cosine (which take revolutions).
__device__ __forceinline__ void phasor_split(float d, float hi, float lo, float &re, float &im) {
#pragma clang fp contract(off)
  const float p = d * hi;
  const float e = fmaf(d, hi, -p);
  const float fr = (p - rintf(p)) + fmaf(d, lo, e);
  re = __builtin_amdgcn_cosf(fr);
  im = -__builtin_amdgcn_sinf(fr);
}

// The lane's taps of one row: the reference's distances and the spectrum
// value at each distance's bin (bin 0 when the distance is out of range; the
// caller zeroes those).  The gathers are issued here and not waited for.
template <int VPT>
__device__ __forceinline__ void interp_taps(int i, const float *__restrict__ ant,
                                            const float2 *__restrict__ meas,
                                            const float *__restrict__ rxo, int R, int S,
                                            float fft_spacing, float inv_fs, float dmax,
                                            int is_ti, const float (&vx)[VPT],
                                            const float (&vy)[VPT], const float (&vz)[VPT],
                                            float (&d)[VPT], float2 (&m)[VPT]) {
  const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
  const int rx = i % R;
  const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
  const bool mono = (ox == 0.0f) & (oy == 0.0f) & (oz == 0.0f);
  const float2 *__restrict__ row = meas + (size_t)i * S;
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    d[k] = interp_dist(vx[k], vy[k], vz[k], px, py, pz, ox, oy, oz, mono, is_ti);
    const bool in = !(d[k] < 0.0f || d[k] > dmax);
    const int b = interp_bin_fast(d[k], fft_spacing, inv_fs);
    // a 32-bit byte offset from the row (bins < S <= 2^28): the load takes
    // the row's address from scalar registers and one VGPR of offset
    m[k] = *(const float2 *)((const char *)row + (uint32_t)(in ? b : 0) * 8u);
  }
}

template <int VPT>
__global__ __launch_bounds__(kBlock) void k_bp_interp(
    float2 *__restrict__ img, float2 *__restrict__ mom,
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    int64_t nvox, const float *__restrict__ ant,
    const float2 *__restrict__ meas, const float *__restrict__ rxo, float il_hi, float il_lo,
    float fft_spacing, float inv_fs, int ny, int nz, int rows, int R, int S, int is_ti) {
  const int64_t base = (int64_t)blockIdx.x * (kBlock * VPT) + threadIdx.x;
  float vx[VPT], vy[VPT], vz[VPT];
  load_voxels<VPT>(xl, yl, zl, vidx, nvox, ny, nz, base, vx, vy, vz);
  settle_voxels<VPT>(vx, vy, vz);
  float Ir[VPT], Ii[VPT], M[VPT][6];
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    Ir[k] = Ii[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) M[k][q] = 0.0f;
  }
  const float dmax = fft_spacing * (float)S;
  // row i's phasors and sums from its taps.  The moment is accumulated as
  // sum_i a_i c_i with the row's vote point a_i = p_i + o_i/2 (wave-uniform,
  // a scalar operand of the FMAs); v * sum_i c_i = v * I is taken off once
  // per voxel at the end, which gives sum_i (a_i - v) c_i.
  auto consume = [&](int i, const float (&d)[VPT], const float2 (&mr)[VPT]) {
    const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
    const int rx = i % R;
    const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
    const float vax = half_off(px, ox), vay = half_off(py, oy), vaz = half_off(pz, oz);
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      float pr, pi;
      phasor_split(d[k], il_hi, il_lo, pr, pi);
      // out-of-range rows read bin 0 and contribute nothing
      const bool in = !(d[k] < 0.0f || d[k] > dmax);
      const float2 m = in ? mr[k] : make_float2(0.0f, 0.0f);
      const float cr = m.x * pr - m.y * pi;
      const float ci = m.x * pi + m.y * pr;
      Ir[k] += cr;
      Ii[k] += ci;
      M[k][0] = fmaf(vax, cr, M[k][0]);
      M[k][1] = fmaf(vax, ci, M[k][1]);
      M[k][2] = fmaf(vay, cr, M[k][2]);
      M[k][3] = fmaf(vay, ci, M[k][3]);
      M[k][4] = fmaf(vaz, cr, M[k][4]);
      M[k][5] = fmaf(vaz, ci, M[k][5]);
    }
  };
  auto taps = [&](int i, float (&d)[VPT], float2 (&m)[VPT]) {
    interp_taps<VPT>(i, ant, meas, rxo, R, S, fft_spacing, inv_fs, dmax, is_ti, vx, vy, vz, d,
                     m);
  };
  for (int i = 0; i < rows; ++i) {
    float d[VPT];
    float2 mr[VPT];
    taps(i, d, mr);
    consume(i, d, mr);
  }
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    M[k][0] = fmaf(-vx[k], Ir[k], M[k][0]);
    M[k][1] = fmaf(-vx[k], Ii[k], M[k][1]);
    M[k][2] = fmaf(-vy[k], Ir[k], M[k][2]);
    M[k][3] = fmaf(-vy[k], Ii[k], M[k][3]);
    M[k][4] = fmaf(-vz[k], Ir[k], M[k][4]);
    M[k][5] = fmaf(-vz[k], Ii[k], M[k][5]);
  }
  store_voxels<VPT>(img, mom, nvox, base, Ir, Ii, M);
}

// --------------------------------------------------------------------------
// Recurrence.  For row i and voxel v the samples' phasors are
//   exp(-j 2pi f_j d/c) = A_b * r^(j - bK),  A_b = exp(-j 2pi f_bK d/c),
//   r = exp(-j 2pi df d/c),
// so the row's block sum  sum_j s_j phi_j = A_b * h_b  with the polynomial
// h_b = sum_t s_{bK+t} r^t evaluated by Horner from the top coefficient down.
// A_b is recomputed from d in FP64 at every block (the re-anchor).  NB blocks
// are evaluated side by side so that each lane carries 2*VPT*NB independent
// FMA chains (a Horner step is two dependent FMAs per component).
// --------------------------------------------------------------------------
template <int VPT>
__device__ __forceinline__ void horner_step(float sr, float si,
                                            const float (&rr)[VPT],
                                            const float (&ri)[VPT],
                                            float (&hr)[VPT],
                                            float (&hi)[VPT]) {
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    const float tr = fmaf(-hi[k], ri[k], fmaf(hr[k], rr[k], sr));
    const float ti = fmaf(hi[k], rr[k], fmaf(hr[k], ri[k], si));
    hr[k] = tr;
    hi[k] = ti;
  }
}

// cr/ci += A(d * fb) * h, with fb the block's first frequency over c
// (revolutions per metre): the FP64 re-anchor.
template <int VPT>
__device__ __forceinline__ void anchor_add(const double (&d)[VPT], double fb,
                                           const float (&hr)[VPT],
                                           const float (&hi)[VPT],
                                           float (&cr)[VPT], float (&ci)[VPT]) {
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    float ar, ai;
    phasor_hw(d[k] * fb, ar, ai);
    cr[k] = fmaf(ar, hr[k], fmaf(-ai, hi[k], cr[k]));
    ci[k] = fmaf(ar, hi[k], fmaf(ai, hr[k], ci[k]));
  }
}

// One full block of K samples starting at sample b0.  S = 1: plain Horner in
// r.  S = 2: the even and odd coefficients as two Horner chains in r^2,
// h = E(r^2) + r O(r^2) -- the same 4 FMAs per sample with twice the
// independent chains per lane.
// h = sum_t blk[t] r^t over KB samples, by Horner from the top coefficient.
template <int KB, int VPT>
__device__ __forceinline__ void block_h(const float2 *__restrict__ blk,
                                        const float (&rr)[VPT], const float (&ri)[VPT],
                                        float (&hr)[VPT], float (&hi)[VPT]) {
  const float2 top = blk[KB - 1];
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    hr[k] = top.x;
    hi[k] = top.y;
  }
#pragma unroll
  for (int t = KB - 2; t >= 0; --t) {
    const float2 s = blk[t];
    horner_step<VPT>(s.x, s.y, rr, ri, hr, hi);
  }
}

// acc = acc * R + h (complex, per voxel): the second Horner level of a
// re-anchor interval longer than 64 samples, R = r^64.
template <int VPT>
__device__ __forceinline__ void outer_step(const float (&Rr)[VPT], const float (&Ri)[VPT],
                                           const float (&hr)[VPT], const float (&hi)[VPT],
                                           float (&ar)[VPT], float (&ai)[VPT]) {
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    const float tr = fmaf(ar[k], Rr[k], fmaf(-ai[k], Ri[k], hr[k]));
    const float ti = fmaf(ar[k], Ri[k], fmaf(ai[k], Rr[k], hi[k]));
    ar[k] = tr;
    ai[k] = ti;
  }
}

// One re-anchor interval of K samples starting at sample b0.  K <= 64: one
// Horner block.  K > 64 (S = 1 only): K / 64 Horner blocks of 64 samples,
// combined from the top one down by a Horner in R = r^64 (R from the FP64
// range reduction of d 64 df / c), so every chain stays 64 steps long while
// the FP64 anchor comes once per K samples.
template <int K, int VPT, int S>
__device__ __forceinline__ void rec_block(const float2 *__restrict__ row, int b0,
                                          double f0c, double dfc,
                                          const double (&d)[VPT],
                                          const float (&rr)[VPT],
                                          const float (&ri)[VPT],
                                          const float (&r2r)[VPT],
                                          const float (&r2i)[VPT],
                                          const float (&Rr)[VPT],
                                          const float (&Ri)[VPT],
                                          float (&cr)[VPT], float (&ci)[VPT]) {
  const float2 *__restrict__ blk = row + b0;
  if constexpr (K > 64) {
    static_assert(S == 1 && K % 64 == 0, "long re-anchor intervals: plain Horner, 64-blocks");
    float ar[VPT], ai[VPT];
    block_h<64, VPT>(blk + (K - 64), rr, ri, ar, ai);
#pragma unroll
    for (int q = K / 64 - 2; q >= 0; --q) {
      float hr[VPT], hi[VPT];
      block_h<64, VPT>(blk + 64 * q, rr, ri, hr, hi);
      outer_step<VPT>(Rr, Ri, hr, hi, ar, ai);
    }
    anchor_add<VPT>(d, fma((double)b0, dfc, f0c), ar, ai, cr, ci);
  } else if constexpr (S == 1) {
    float hr[VPT], hi[VPT];
    block_h<K, VPT>(blk, rr, ri, hr, hi);
    anchor_add<VPT>(d, fma((double)b0, dfc, f0c), hr, hi, cr, ci);
  } else {
    static_assert(K % 2 == 0, "split Horner needs an even block");
    float er[VPT], ei[VPT], orr[VPT], oi[VPT];
    const float2 te = blk[K - 2], to = blk[K - 1];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      er[k] = te.x;
      ei[k] = te.y;
      orr[k] = to.x;
      oi[k] = to.y;
    }
#pragma unroll
    for (int t = K / 2 - 2; t >= 0; --t) {
      const float2 se = blk[2 * t], so = blk[2 * t + 1];
      horner_step<VPT>(se.x, se.y, r2r, r2i, er, ei);
      horner_step<VPT>(so.x, so.y, r2r, r2i, orr, oi);
    }
    float hr[VPT], hi[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      hr[k] = fmaf(rr[k], orr[k], fmaf(-ri[k], oi[k], er[k]));
      hi[k] = fmaf(rr[k], oi[k], fmaf(ri[k], orr[k], ei[k]));
    }
    anchor_add<VPT>(d, fma((double)b0, dfc, f0c), hr, hi, cr, ci);
  }
}

// Work split (RecPlan): workgroups [0, nmain) take voxel block b = blockIdx.x
// and every row, and write img/mom.  The others take tail voxel block
// nmain + (t mod ntail) and row segment s = t / ntail (t = blockIdx.x - nmain),
// rows [s rps, min((s + 1) rps, rows)), and write their partial sums to
// pimg/pmom + s * (tail voxels), indexed from the first tail voxel; the
// segments are added afterwards in a fixed order.
// Waves per SIMD at least 7 (at most 73 VGPRs) for the default shapes,
// VPT x S <= 2; at least 4 for VPT = 2 with S = 2 and 3 for VPT = 4, which
// need more registers.  With the row loop in scalar registers the compiler
// otherwise takes 75 VGPRs (six waves) and spills 45 SGPRs into VGPR lanes
// at K = 64, VPT = 2; with the cap it fits in 70 VGPRs and 84 SGPRs without
// spills (measured at 128x128x96 voxels, 3600 rows, 256 samples: 109.8 ms
// against 111.6 without the cap and 111.9-112.1 with the row loop on the
// vector ALU, whose results the capped kernel reproduces bit for bit).
template <int K, int VPT, int S>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(VPT >= 4 ? 3 : VPT * S <= 2 ? 7 : 4, 8))) void k_bp_rec(
    float2 *__restrict__ img, float2 *__restrict__ mom,
    float2 *__restrict__ pimg, float2 *__restrict__ pmom,
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    int64_t nvox, const float *__restrict__ ant,
    const float2 *__restrict__ meas, const float *__restrict__ rxo,
    double f0c, double dfc, int ny, int nz, int rows, int R, int F,
    int is_ti, int64_t nmain, int64_t ntail, int rps) {
  // The block's voxel block and row segment, and from them the row range and
  // the first row's receiver, are wave-uniform: readfirstlane keeps them (and
  // the row loop with its address arithmetic) in scalar registers, where the
  // integer division (expanded on the vector ALU) would otherwise leave them.
  // The grid has fewer than 2^32 blocks, so 32-bit arithmetic suffices.
  const unsigned bid = blockIdx.x, nm = (unsigned)nmain;
  unsigned vb = bid;
  int seg = -1;
  if (bid >= nm) {
    const unsigned t = bid - nm, nt = (unsigned)ntail;
    const unsigned sq = t / nt;
    seg = (int)sq;
    vb = nm + (t - sq * nt);
  }
  seg = __builtin_amdgcn_readfirstlane(seg);
  vb = __builtin_amdgcn_readfirstlane(vb);
  const int64_t base = (int64_t)vb * (kBlock * VPT) + threadIdx.x;
  float vx[VPT], vy[VPT], vz[VPT];
  load_voxels<VPT>(xl, yl, zl, vidx, nvox, ny, nz, base, vx, vy, vz);
  float Ir[VPT], Ii[VPT], Cr[VPT], Ci[VPT], M[VPT][6];
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    Ir[k] = Ii[k] = Cr[k] = Ci[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) M[k][q] = 0.0f;
  }
  const int nb = F / K;
  const int tail = F - nb * K;

  const int row_lo = seg < 0 ? 0 : seg * rps;
  const int row_hi = seg < 0 ? rows : min(rows, row_lo + rps);
  // d * 4 df/c equals 4 d * df/c exactly (a power-of-two scale), and adding
  // 0.0 leaves d unchanged: uniform constants instead of per-voxel selects.
  const double dfc4 = 4.0 * dfc, dfc64 = 64.0 * dfc, d_off = is_ti ? 0.15 : 0.0;
  int rx = __builtin_amdgcn_readfirstlane(row_lo % R);
  for (int i = row_lo; i < row_hi; ++i, rx = rx + 1 == R ? 0 : rx + 1) {
    const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
    const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
    const bool mono = (ox == 0.0f) & (oy == 0.0f) & (oz == 0.0f);
    const float2 *__restrict__ row = meas + (size_t)i * F;

    double d[VPT];
    float rr[VPT], ri[VPT], r2r[VPT], r2i[VPT], Rr[VPT], Ri[VPT], cr[VPT], ci[VPT];
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      d[k] = roundtrip_dist64(vx[k], vy[k], vz[k], px, py, pz, ox, oy, oz, mono) + d_off;
      phasor_quarter(d[k] * dfc4, rr[k], ri[k]);
      r2r[k] = fmaf(rr[k], rr[k], -ri[k] * ri[k]);
      r2i[k] = 2.0f * rr[k] * ri[k];
      if constexpr (K > 64) phasor_hw(d[k] * dfc64, Rr[k], Ri[k]);
      else Rr[k] = Ri[k] = 0.0f;
      cr[k] = ci[k] = 0.0f;
    }

    for (int b = 0; b < nb; ++b)
      rec_block<K, VPT, S>(row, b * K, f0c, dfc, d, rr, ri, r2r, r2i, Rr, Ri, cr, ci);
    if (K > 64 && tail) {
      // the last, partial interval: its 64-blocks and the remainder above
      // them, combined from the top down as in rec_block
      const int g0 = nb * K, nfull = tail / 64, rem = tail - 64 * nfull;
      float ar[VPT], ai[VPT];
      int q = nfull - 1;
      if (rem) {
        const float2 *__restrict__ blk = row + g0 + 64 * nfull;
        const float2 top = blk[rem - 1];
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
          ar[k] = top.x;
          ai[k] = top.y;
        }
        for (int t = rem - 2; t >= 0; --t) {
          const float2 s = blk[t];
          horner_step<VPT>(s.x, s.y, rr, ri, ar, ai);
        }
      } else {
        block_h<64, VPT>(row + g0 + 64 * q, rr, ri, ar, ai);
        --q;
      }
      for (; q >= 0; --q) {
        float hr[VPT], hi[VPT];
        block_h<64, VPT>(row + g0 + 64 * q, rr, ri, hr, hi);
        outer_step<VPT>(Rr, Ri, hr, hi, ar, ai);
      }
      anchor_add<VPT>(d, fma((double)g0, dfc, f0c), ar, ai, cr, ci);
    } else if (tail) {
      const float2 *__restrict__ blk = row + nb * K;
      float hr[VPT], hi[VPT];
      {
        const float2 top = blk[tail - 1];
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
          hr[k] = top.x;
          hi[k] = top.y;
        }
      }
      for (int t = tail - 2; t >= 0; --t) {
        const float2 s = blk[t];
        horner_step<VPT>(s.x, s.y, rr, ri, hr, hi);
      }
      anchor_add<VPT>(d, fma((double)(nb * K), dfc, f0c), hr, hi, cr, ci);
    }

    const float vax = px + ox / 2.0f, vay = py + oy / 2.0f, vaz = pz + oz / 2.0f;
#pragma unroll
    for (int k = 0; k < VPT; ++k)
      add_row(cr[k], ci[k], vax - vx[k], vay - vy[k], vaz - vz[k], Ir[k],
              Ii[k], Cr[k], Ci[k], M[k]);
  }
  if (seg < 0) {
    store_voxels<VPT>(img, mom, nvox, base, Ir, Ii, M);
  } else {
    const int64_t off = nmain * (kBlock * VPT), ntv = nvox - off;
    store_voxels<VPT>(pimg + (size_t)seg * ntv, pmom ? pmom + (size_t)seg * ntv * 3 : nullptr,
                      nvox, base, Ir, Ii, M, off);
  }
}

// Fixed-order FP64 sum of the row segments' partial sums.
__global__ __launch_bounds__(kBlock) void k_rs_reduce(
    float2 *__restrict__ img, float2 *__restrict__ mom,
    const float2 *__restrict__ pimg, const float2 *__restrict__ pmom, int64_t n,
    int rs) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < rs; ++q) {
    const float2 v = pimg[(size_t)q * n + t];
    a += v.x;
    b += v.y;
  }
  img[t] = make_float2((float)a, (float)b);
  if (mom) {
    for (int k = 0; k < 3; ++k) {
      double c = 0.0, e = 0.0;
      for (int q = 0; q < rs; ++q) {
        const float2 v = pmom[((size_t)q * n + t) * 3 + k];
        c += v.x;
        e += v.y;
      }
      mom[3 * t + k] = make_float2((float)c, (float)e);
    }
  }
}

int env_int(const char *name, int dflt) {
  const char *s = getenv(name);
  return (s && *s) ? atoi(s) : dflt;
}

struct RecArgs {
  float2 *img, *mom;
  const float *xl, *yl, *zl;
  const int32_t *vidx;
  int64_t nvox;
  const float *ant;
  const float2 *meas;
  const float *rxo;
  double f0c, dfc;
  int ny, nz, rows, R, F, is_ti;
  void *ws;        // row-segment partials (may be null)
  size_t ws_bytes;
};

// Most row segments a launch may use: never fewer than 32 rows per segment.
int rs_cap(int rows) {
  int cap = rows / 32;
  if (cap > 32) cap = 32;
  if (cap < 1) cap = 1;
  return cap;
}

// Resident workgroups of a kernel on the current device (occupancy x CUs).
// Without a device (the workspace query may run on a host with no GPU) it
// assumes MI355X and the recurrence kernel's measured occupancy: 256 CUs, 8
// workgroups of 256 lanes per CU.
template <typename KernelT>
int64_t resident_slots(KernelT kernel) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 8;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  (void)hipGetLastError();
  return (int64_t)per_cu * cus;
}

// Work split of one recurrence launch (see k_bp_rec).  When the voxel
// blocks give fewer than eight rounds of resident workgroups, every block is
// split over the fewest row segments that give eight (at least 32 rows per
// segment, at most 32 segments), and the segments' partial sums pass
// through the workspace.  Measured on MI355X at 128x128x96 voxels and 3600
// rows (3072 blocks, 1792 resident): every block in 4-9 segments 111-113 ms,
// only the last partial round split 114.5 ms, no split 121.9 ms.  Splitting
// every block also keeps each workgroup's window of measurement rows (720
// rows, 1.5 MB at five segments) inside an XCD's 4 MB L2: FETCH_SIZE read
// 59 MB per launch, one pass over the rows per XCD, against 378 / 489 MB for
// the other two.  MMN_REC_RS = n forces n segments.
// The kernel's mapping also allows whole blocks and split blocks in one
// launch; the plan uses one or the other.
struct RecPlan {
  int64_t nmain = 0, ntail = 0;  // voxel blocks taking all rows / split
  int nseg = 1, rps = 0;
  int64_t tail_vox = 0;          // voxels of the split blocks
  size_t ws = 0;                 // workspace bytes of the partial sums
};

RecPlan plan_rec(int64_t nvox, int rows, bool moment, int vpt, int64_t slots,
                 size_t ws_limit) {
  const int64_t per_blk = (int64_t)kBlock * vpt;
  const int64_t nblk = (nvox + per_blk - 1) / per_blk;
  RecPlan p;
  p.nmain = nblk;
  p.rps = rows;
  int nseg = 1;
  int64_t ntail = 0;
  const int forced = env_int("MMN_REC_RS", 0);
  if (forced > 0) {
    nseg = forced;
    ntail = nblk;
  } else if (nblk < 8 * slots) {
    ntail = nblk;
    while (ntail * nseg < 8 * slots && nseg < 64) ++nseg;
  }
  nseg = std::min(nseg, rs_cap(rows));
  for (; nseg > 1; --nseg) {
    const int rps = (rows + nseg - 1) / nseg;
    const int used = (rows + rps - 1) / rps;
    const int64_t tv = nvox - (nblk - ntail) * per_blk;
    const size_t bytes = (size_t)used * (size_t)tv * sizeof(float2) * (moment ? 4 : 1);
    if (bytes > ws_limit) continue;
    if (used < 2) break;
    p.nmain = nblk - ntail;
    p.ntail = ntail;
    p.nseg = used;
    p.rps = rps;
    p.tail_vox = tv;
    p.ws = bytes;
    break;
  }
  return p;
}

template <int K, int VPT, int S>
void launch_rec(hipStream_t st, const RecArgs &a) {
  const int64_t slots = resident_slots(k_bp_rec<K, VPT, S>);
  const RecPlan p = plan_rec(a.nvox, a.rows, a.mom != nullptr, VPT, slots,
                             a.ws ? a.ws_bytes : 0);
  float2 *pimg = nullptr, *pmom = nullptr;
  if (p.nseg > 1) {
    pimg = (float2 *)a.ws;
    pmom = a.mom ? pimg + (size_t)p.nseg * p.tail_vox : nullptr;
  }
  const int64_t nwg = p.nmain + p.ntail * p.nseg;
  hipLaunchKernelGGL((k_bp_rec<K, VPT, S>), dim3((unsigned)nwg), dim3(kBlock), 0, st, a.img,
                     a.mom, pimg, pmom, a.xl, a.yl, a.zl, a.vidx, a.nvox, a.ant, a.meas,
                     a.rxo, a.f0c, a.dfc, a.ny, a.nz, a.rows, a.R, a.F, a.is_ti, p.nmain,
                     p.ntail, p.rps);
  if (p.nseg > 1) {
    const int64_t off = a.nvox - p.tail_vox;
    mmn_rs_reduce(st, a.img + off, a.mom ? a.mom + 3 * off : nullptr, pimg, pmom, p.tail_vox,
                  p.nseg);
  }
}

template <int VPT, int S>
int dispatch_k(int K, hipStream_t st, const RecArgs &a) {
  switch (K) {
    case 8:  launch_rec<8, VPT, S>(st, a); return 0;
    case 16: launch_rec<16, VPT, S>(st, a); return 0;
    case 32: launch_rec<32, VPT, S>(st, a); return 0;
    case 64: launch_rec<64, VPT, S>(st, a); return 0;
    default: break;
  }
  if constexpr (S == 1) {  // intervals longer than one Horner block
    if (K == 128) { launch_rec<128, VPT, S>(st, a); return 0; }
    if (K == 256) { launch_rec<256, VPT, S>(st, a); return 0; }
  }
  return -1;
}

// Lane shape: VPT voxels per lane, S-way split Horner.  The defaults (2, 1)
// were chosen by measurement on MI355X (config 2, K = 64): VPT 2 beats 1 by
// ~8 %, and VPT 4 (128 VGPRs, four waves per SIMD, half the scalar-load waits
// per FMA) is 0.5 % slower; the even/odd split and two blocks side by side give no gain (the
// latter spills SGPRs); staging rows in LDS with a double-buffered tile, or
// streaming chunks with explicit prefetch, measured 15-50 % slower (their
// extra VGPRs cost occupancy).  MMN_REC_VPT / MMN_REC_SPLIT override them.
int dispatch_rec(int K, hipStream_t st, const RecArgs &a) {
  const int vpt = env_int("MMN_REC_VPT", 2);
  const int split = env_int("MMN_REC_SPLIT", 1);
  if (vpt == 4) return dispatch_k<4, 1>(K, st, a);
  if (vpt == 1 && split == 1) return dispatch_k<1, 1>(K, st, a);
  if (vpt == 1) return dispatch_k<1, 2>(K, st, a);
  if (split == 1) return dispatch_k<2, 1>(K, st, a);
  return dispatch_k<2, 2>(K, st, a);
}

// Workspace the recurrence launch uses, for the lane shape dispatch_rec
// picks and the largest over the re-anchor intervals (their occupancy may
// differ).
template <int VPT, int S>
size_t ws_for_shape(int64_t n, int rows, bool mom) {
  const size_t lim = (size_t)-1;
  size_t m = plan_rec(n, rows, mom, VPT, resident_slots(k_bp_rec<8, VPT, S>), lim).ws;
  m = std::max(m, plan_rec(n, rows, mom, VPT, resident_slots(k_bp_rec<16, VPT, S>), lim).ws);
  m = std::max(m, plan_rec(n, rows, mom, VPT, resident_slots(k_bp_rec<32, VPT, S>), lim).ws);
  m = std::max(m, plan_rec(n, rows, mom, VPT, resident_slots(k_bp_rec<64, VPT, S>), lim).ws);
  if constexpr (S == 1) {
    m = std::max(m, plan_rec(n, rows, mom, VPT, resident_slots(k_bp_rec<128, VPT, S>), lim).ws);
    m = std::max(m, plan_rec(n, rows, mom, VPT, resident_slots(k_bp_rec<256, VPT, S>), lim).ws);
  }
  return m;
}

size_t ws_for_rec(int64_t n, int rows, bool mom) {
  const int vpt = env_int("MMN_REC_VPT", 2);
  const int split = env_int("MMN_REC_SPLIT", 1);
  if (vpt == 4) return ws_for_shape<4, 1>(n, rows, mom);
  if (vpt == 1 && split == 1) return ws_for_shape<1, 1>(n, rows, mom);
  if (vpt == 1) return ws_for_shape<1, 2>(n, rows, mom);
  if (split == 1) return ws_for_shape<2, 1>(n, rows, mom);
  return ws_for_shape<2, 2>(n, rows, mom);
}

}  // namespace

void mmn_rs_reduce(hipStream_t st, float2 *img, float2 *mom, const float2 *pimg,
                   const float2 *pmom, int64_t n, int rs) {
  hipLaunchKernelGGL(k_rs_reduce, dim3(mmn_blocks(n, kBlock)), dim3(kBlock), 0, st, img, mom,
                     pimg, pmom, n, rs);
}

extern "C" int mmn_bp_interp(float *img, float *moment, const float *x_locs,
                             const float *y_locs, const float *z_locs,
                             const int32_t *vox_idx, int64_t n_vox, const float *ant,
                             const float *meas, const float *rx_off, float wavelength,
                             float fft_spacing, int nx, int ny, int nz, int num_rows,
                             int num_rx, int samples, int is_ti_radar, void *stream) {
  MMN_REQUIRE(img && x_locs && y_locs && z_locs && ant && meas && rx_off, MMN_ERR_NULL,
              "mmn_bp_interp: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 && samples > 0 &&
                  wavelength > 0.0f && fft_spacing > 0.0f,
              MMN_ERR_ARG, "mmn_bp_interp: sizes and scalars must be positive");
  int64_t nvox = (int64_t)nx * ny * nz;
  if (vox_idx) {
    MMN_REQUIRE(n_vox >= 0, MMN_ERR_ARG, "mmn_bp_interp: n_vox < 0");
    nvox = n_vox;
  }
  if (nvox == 0) return 0;
  MMN_REQUIRE(samples <= (1 << 28), MMN_ERR_ARG, "mmn_bp_interp: more than 2^28 samples a row");
  const double il = 1.0 / (double)wavelength;
  const float il_hi = (float)il, il_lo = (float)(il - (double)il_hi);
  hipLaunchKernelGGL(k_bp_interp<2>, dim3(mmn_blocks(nvox, kBlock * 2)), dim3(kBlock), 0,
                     (hipStream_t)stream, (float2 *)img, (float2 *)moment, x_locs, y_locs,
                     z_locs, vox_idx, nvox, ant, (const float2 *)meas, rx_off, il_hi, il_lo,
                     fft_spacing, 1.0f / fft_spacing, ny, nz, num_rows, num_rx, samples,
                     is_ti_radar);
  return mmn_check_launch("mmn_bp_interp");
}

extern "C" int mmn_sqrt_rn_check(uint64_t first, uint64_t count, unsigned long long *result,
                                 void *stream) {
  MMN_REQUIRE(result, MMN_ERR_NULL, "mmn_sqrt_rn_check: NULL result");
  MMN_REQUIRE(count <= (1ull << 32) && first + count <= (1ull << 32), MMN_ERR_ARG,
              "mmn_sqrt_rn_check: the range must lie within the 2^32 float bit patterns");
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_sqrt_rn_check, dim3(8192), dim3(kBlock), 0, (hipStream_t)stream, result,
                     first, count);
  return mmn_check_launch("mmn_sqrt_rn_check");
}

extern "C" int mmn_cuda_image(float *img, const float *x_locs,
                              const float *y_locs, const float *z_locs,
                              const float *ant, const float *meas,
                              const float *rx_off, float slope,
                              float wavelength, float fft_spacing, int nx,
                              int ny, int nz, int num_rows, int num_rx,
                              int samples, int is_ti_radar,
                              int use_interpolated, void *stream) {
  MMN_REQUIRE(img && x_locs && y_locs && z_locs && ant && meas && rx_off,
              MMN_ERR_NULL, "mmn_cuda_image: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 &&
                  samples > 0,
              MMN_ERR_ARG,
              "mmn_cuda_image: sizes must be positive (nx=%d ny=%d nz=%d "
              "rows=%d rx=%d samples=%d)",
              nx, ny, nz, num_rows, num_rx, samples);
  const int64_t nvox = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(k_image_mirror, dim3(mmn_blocks(nvox, kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, (float2 *)img,
                     x_locs, y_locs, z_locs, ant, (const float2 *)meas, rx_off,
                     slope, wavelength, fft_spacing, ny, nz, nvox, num_rows,
                     num_rx, samples, is_ti_radar, use_interpolated);
  return mmn_check_launch("mmn_cuda_image");
}

extern "C" size_t mmn_bp_forward_workspace_size(int64_t n_vox, int num_rows,
                                                int with_moment) {
  if (n_vox <= 0 || num_rows <= 0) return 0;
  return ws_for_rec(n_vox, num_rows, with_moment != 0);
}

extern "C" int mmn_bp_forward(float *img, float *moment, const float *x_locs,
                              const float *y_locs, const float *z_locs,
                              const int32_t *vox_idx, int64_t n_vox,
                              const float *ant, const float *meas,
                              const float *rx_off, double f0, double df,
                              int nx, int ny, int nz, int num_rows,
                              int num_rx, int samples, int is_ti_radar,
                              int variant, int K, void *workspace,
                              size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(img && x_locs && y_locs && z_locs && ant && meas && rx_off,
              MMN_ERR_NULL, "mmn_bp_forward: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 &&
                  samples > 0,
              MMN_ERR_ARG, "mmn_bp_forward: sizes must be positive");
  MMN_REQUIRE(variant == MMN_BP_PLAIN || variant == MMN_BP_REC ||
                  variant == MMN_BP_PLAIN_FAST,
              MMN_ERR_ARG,
              "mmn_bp_forward: unknown variant %d", variant);
  if (K == 0) K = 64;
  MMN_REQUIRE(variant != MMN_BP_REC || K == 8 || K == 16 || K == 32 || K == 64 ||
                  K == 128 || K == 256,
              MMN_ERR_ARG, "mmn_bp_forward: K must be 8, 16, 32, 64, 128 or 256 (got %d)",
              K);
  const int64_t ngrid = (int64_t)nx * ny * nz;
  int64_t nvox = ngrid;
  if (vox_idx) {
    MMN_REQUIRE(n_vox >= 0, MMN_ERR_ARG, "mmn_bp_forward: n_vox < 0");
    nvox = n_vox;
  }
  if (nvox == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  float2 *im = (float2 *)img;
  float2 *mo = (float2 *)moment;
  const float2 *ms = (const float2 *)meas;
  const double inv_c = 1.0 / MMN_C0;
  if (variant == MMN_BP_PLAIN || variant == MMN_BP_PLAIN_FAST) {
    const float k0 = (float)(2.0 * kPi * f0 * inv_c);
    const float dk = (float)(2.0 * kPi * df * inv_c);
    const float u0 = (float)(f0 * inv_c), du = (float)(df * inv_c);
    const bool fast = variant == MMN_BP_PLAIN_FAST;
    // Voxels per lane (MMN_PLAIN_VPT = 1, 2 or 4 overrides): measured at
    // config 2 with the samples in groups of 8, the library-sincosf kernel
    // 881 / 913 / 1152 ms at 1 / 2 / 4 and the hardware sin/cos kernel
    // 453 / 430 / 432 ms (3430 and 457 ms with one load and wait per sample).
    static const int pvpt_env = env_int("MMN_PLAIN_VPT", 0);
    const int pvpt = pvpt_env == 1 || pvpt_env == 2 || pvpt_env == 4 ? pvpt_env
                                                                      : (fast ? 2 : 1);
    auto kern = fast ? (pvpt == 4 ? k_bp_plain<4, true> : pvpt == 2 ? k_bp_plain<2, true>
                                                                    : k_bp_plain<1, true>)
                     : (pvpt == 4 ? k_bp_plain<4, false> : pvpt == 2 ? k_bp_plain<2, false>
                                                                     : k_bp_plain<1, false>);
    hipLaunchKernelGGL(kern, dim3(mmn_blocks(nvox, kBlock * pvpt)), dim3(kBlock), 0, st, im, mo,
                       x_locs, y_locs, z_locs, vox_idx, nvox, ant, ms, rx_off, k0, dk, u0, du,
                       ny, nz, num_rows, num_rx, samples, is_ti_radar);
    return mmn_check_launch("mmn_bp_forward(plain)");
  }
  RecArgs ra{im, mo, x_locs, y_locs, z_locs, vox_idx, nvox, ant, ms, rx_off,
             f0 * inv_c, df * inv_c, ny, nz, num_rows, num_rx,
             samples, is_ti_radar, workspace, workspace_bytes};
  const int rc = dispatch_rec(K, st, ra);
  MMN_REQUIRE(rc == 0, MMN_ERR_ARG, "mmn_bp_forward: bad K %d", K);
  return mmn_check_launch("mmn_bp_forward(rec)");
}
