// Range-profile backprojection: the full-frequency sum of mmn_bp_forward,
// evaluated per (row, voxel) from a per-row table instead of per sample.
//
// For row i write f_t = f0 + t df and centre the band at t_c = (F - 1) / 2:
//   P_i(d) = sum_t s_it exp(-j 2 pi f_t d / c) = exp(-j 2 pi f_c d / c) Q_i(d),
//   Q_i(d) = sum_t s_it exp(-j 2 pi (t - t_c) df d / c),   f_c = f0 + t_c df.
// Q_i is a trigonometric polynomial in d with no component faster than
// (F - 1) df / (2c) cycles per metre: it is band-limited, smooth on the scale
// of the range resolution c / ((F - 1) df).  Sampled at h = c / (64 (F - 1) df),
// 64 points per resolution cell, the cubic through the four samples around d
// reproduces Q_i to a few 1e-8 of sum_t |s_it| (the Lagrange remainder of a
// tone at pi/64 rad per sample is at most 0.0234 (pi/64)^4 = 1.4e-7 of its
// amplitude, at mid-interval).  So:
//   k_prof_rows   per row, the table start (a lower bound of the row's
//                 round-trip distance over the grid's bounding box, less h)
//                 and the FP64 constants of the per-voxel evaluation;
//   k_prof_table  Q_i at M points h apart, by Horner in exp(-j 2 pi df d / c)
//                 re-anchored every 64 samples by an FP64-reduced phasor (the
//                 recurrence kernel's arithmetic);
//   k_prof_coef   each interval's cubic as four complex coefficients (32 B),
//                 times the carrier exp(-j 2 pi f_c d_m / c) at the interval
//                 start d_m (FP64 phase);
//   k_bp_prof     per (row, voxel): the FP64 round-trip distance, interval
//                 and fraction, one 32-byte coefficient fetch, the cubic
//                 (6 FMA), the carrier's remainder past the interval start
//                 (under half a turn: FP32 phase, hardware sine/cosine), the
//                 vote moment as the recurrence kernel forms it, and the row
//                 sum -- plain FP32 within blocks of 8 rows, Kahan-compensated
//                 across blocks.  The TI offset and a monostatic row's factor
//                 2 are folded into per-row constants.
// A round trip over the box spans at most twice its diagonal, so M =
// ceil(2 diag / h) + 6 points cover every voxel with the cubic's neighbours.
//
// Full-grid launches map each wave to a 4x4x4 brick of voxels (a 256-lane
// workgroup to 8x8x4, twice that in z with two voxels per lane): the
// brick's distances to a row span a few centimetres, so a wave's coefficient
// fetches fall on a few cache lines of the row's table.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hipcub/hipcub.hpp>

#include "mmn_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kOver = 64;     // table points per range-resolution cell
constexpr int kHornerK = 64;  // samples per FP64 re-anchor in the table
constexpr int kMaxSeg = 8;    // most row segments of one launch
constexpr int kMaxM = 1 << 22;
constexpr int kRowBlock = 8;  // rows summed in plain FP32 between Kahan steps

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct ProfGeom {
  double h;
  int M;
};

// Table spacing and length for F samples from f0, df apart, over a box
// (host).  h is also capped at half the centre wavelength, which keeps the
// carrier's remainder within an interval under half a turn.
bool prof_geom(int F, double f0, double df, const double *bb, ProfGeom *g) {
  const double ex = bb[3] - bb[0], ey = bb[4] - bb[1], ez = bb[5] - bb[2];
  if (!(ex >= 0.0 && ey >= 0.0 && ez >= 0.0) || !isfinite(ex + ey + ez)) return false;
  if (!isfinite(f0) || !isfinite(df)) return false;
  const double diag = sqrt(ex * ex + ey * ey + ez * ez);
  const double band = (double)(F - 1) * fabs(df);
  const double fc = fabs(f0 + 0.5 * (double)(F - 1) * df);
  // with one sample or df = 0, Q_i is constant and any spacing is exact
  double h = band > 0.0 ? MMN_C0 / (kOver * band) : (2.0 * diag + 1e-3) / 8.0;
  if (fc > 0.0 && h > 0.5 * MMN_C0 / fc) h = 0.5 * MMN_C0 / fc;
  g->h = h;
  const double m = ceil(2.0 * diag / g->h) + 6.0;
  if (!(m <= (double)kMaxM)) return false;
  g->M = (int)m;
  return true;
}

int seg_cap(int rows) {
  int cap = rows / 32;
  if (cap > kMaxSeg) cap = kMaxSeg;
  return cap < 1 ? 1 : cap;
}

// Per-row constants of k_bp_prof, 12 doubles: u offset, |p|^2, -2p,
// |p+o|^2, -2(p+o), the distance-to-u scale (doubled for a monostatic row,
// whose round trip is twice one leg), pad.
constexpr int kRowC = 16;

struct ProfWs {
  double *dlo;   // [A]
  double *rowc;  // [A][kRowC]
  float2 *tab;   // [A][M]
  float4 *coef;  // [A][M][2]
  float2 *part;  // [segments][n] images, then [segments][n][3] moments
};

size_t prof_layout(int A, int M, int64_t n, bool moment, ProfWs *w, char *base) {
  const int segs = seg_cap(A);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  char *d = take(sizeof(double) * (size_t)A);
  char *rc = take(sizeof(double) * kRowC * (size_t)A);
  char *t = take(sizeof(float2) * (size_t)A * M);
  char *c = take(sizeof(float4) * 2 * (size_t)A * M);
  char *p = segs > 1 ? take(sizeof(float2) * (size_t)segs * n * (moment ? 4 : 1)) : nullptr;
  if (w) {
    w->dlo = (double *)d;
    w->rowc = (double *)rc;
    w->tab = (float2 *)t;
    w->coef = (float4 *)c;
    w->part = (float2 *)p;
  }
  return off;
}

// Distance from a point to the box [lo, hi] (0 inside).
__device__ __forceinline__ double box_dist(double px, double py, double pz, const double *bb) {
  const double qx = fmin(fmax(px, bb[0]), bb[3]) - px;
  const double qy = fmin(fmax(py, bb[1]), bb[4]) - py;
  const double qz = fmin(fmax(pz, bb[2]), bb[5]) - pz;
  return sqrt(fma(qz, qz, fma(qy, qy, qx * qx)));
}

struct Box {
  double b[6];
};

__global__ __launch_bounds__(kBlock) void k_prof_rows(double *__restrict__ dlo,
                                                      double *__restrict__ rowc,
                                                      const float *__restrict__ ant,
                                                      const float *__restrict__ rxo, Box box,
                                                      int A, int R, int is_ti, double h) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= A) return;
  const double px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
  const int rx = i % R;
  const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
  const bool mono = (ox == 0.0f) & (oy == 0.0f) & (oz == 0.0f);
  const double qx = px + (double)ox, qy = py + (double)oy, qz = pz + (double)oz;
  const double ti = is_ti ? 0.15 : 0.0;
  const double lo = box_dist(px, py, pz, box.b) + box_dist(qx, qy, qz, box.b) + ti - h;
  dlo[i] = lo;  // table point k sits at lo + k h; voxels fall at k >= 1
  const double inv_h = 1.0 / h, two = mono ? 2.0 : 1.0;
  double *__restrict__ c = rowc + (size_t)kRowC * i;
  // u = (d_geom + ti - lo) / h with d_geom = two * (the legs as evaluated)
  c[0] = (ti - lo) * inv_h;
  c[1] = fma(pz, pz, fma(py, py, px * px));
  c[2] = -2.0 * px;
  c[3] = -2.0 * py;
  c[4] = -2.0 * pz;
  c[5] = fma(qz, qz, fma(qy, qy, qx * qx));
  c[6] = -2.0 * qx;
  c[7] = -2.0 * qy;
  c[8] = -2.0 * qz;
  c[9] = two * inv_h;
  // the row's flag and the moment's point p + o/2 (FP32, exact in FP64), so
  // that the voxel loop reads one record per row
  c[10] = mono ? 1.0 : 0.0;
  // the point's three floats as raw bits in c[11], c[12] (no conversions in
  // the voxel loop)
  const float2 va01 = make_float2(__fadd_rn(ant[3 * i + 0], __fmul_rn(ox, 0.5f)),
                                  __fadd_rn(ant[3 * i + 1], __fmul_rn(oy, 0.5f)));
  const float2 va2 = make_float2(__fadd_rn(ant[3 * i + 2], __fmul_rn(oz, 0.5f)), 0.0f);
  c[11] = __builtin_bit_cast(double, va01);
  c[12] = __builtin_bit_cast(double, va2);
  c[13] = 0.0;
  c[14] = 0.0;
  c[15] = 0.0;
}

// One workgroup row of lanes per table row (blockIdx.y): the row's samples
// are wave-uniform and arrive through scalar loads.
// With kscale the samples are weighted by k_j = 2 pi f_j / c = k0 + j dk
// (the derivative profile the backward's position term reads).
__global__ __launch_bounds__(kBlock) void k_prof_table(
    float2 *__restrict__ tab, const float2 *__restrict__ meas,
    const double *__restrict__ dlo, double h, double dfc, double tc, int M, int F,
    int kscale, float k0, float dk) {
  const int i = blockIdx.y;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  const double d = dlo[i] + (double)min(k, M - 1) * h;
  float rr, ri;
  phasor_quarter(4.0 * d * dfc, rr, ri);  // exp(-j 2 pi df d / c)
  const float2 *__restrict__ row = meas + (size_t)i * F;
  float cr = 0.0f, ci = 0.0f;
  for (int b0 = 0; b0 < F; b0 += kHornerK) {
    const int len = min(kHornerK, F - b0);
    auto sample = [&](int j) {
      float2 v = row[j];
      if (kscale) {
        const float kj = fmaf((float)j, dk, k0);
        v.x *= kj;
        v.y *= kj;
      }
      return v;
    };
    const float2 top = sample(b0 + len - 1);
    float hr = top.x, hi = top.y;
    for (int t = len - 2; t >= 0; --t) {
      const float2 s = sample(b0 + t);
      const float tr = fmaf(-hi, ri, fmaf(hr, rr, s.x));
      const float ti = fmaf(hi, rr, fmaf(hr, ri, s.y));
      hr = tr;
      hi = ti;
    }
    float ar, ai;
    phasor_hw(d * (((double)b0 - tc) * dfc), ar, ai);
    cr = fmaf(ar, hr, fmaf(-ai, hi, cr));
    ci = fmaf(ar, hi, fmaf(ai, hr, ci));
  }
  if (k < M) tab[(size_t)i * M + k] = make_float2(cr, ci);
}

// The cubic through table points m-1 .. m+2 as a polynomial in the fraction
// f in [0, 1) of interval [m, m+1]:
//   Q(f) = q + c1 f + c2 f^2 + c3 f^3,  c1 = -p/3 - q/2 + r - s/6,
//   c2 = (p + r)/2 - q,  c3 = (s - p)/6 + (q - r)/2,
// times the carrier at the interval start, A_m = exp(-j 2 pi f_c d_m / c)
// with d_m = dlo + m h (FP64 phase, FP64 sine/cosine): the voxel then needs
// only the carrier's remainder exp(-j 2 pi (f_c h / c) f), under half a turn
// per interval.  Stored as A_m (q, c1) and A_m (c2, c3); intervals without
// four neighbours hold 0.
__global__ __launch_bounds__(kBlock) void k_prof_coef(float4 *__restrict__ coef,
                                                      const float2 *__restrict__ tab,
                                                      const double *__restrict__ dlo, double h,
                                                      double fcc, int M, int64_t AM) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= AM) return;
  const int m = (int)(t % M);
  float4 lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), hi = lo;
  if (m >= 1 && m <= M - 3) {
    const float2 pf = tab[t - 1], qf = tab[t], rf = tab[t + 1], sf = tab[t + 2];
    const double p[2] = {pf.x, pf.y}, q[2] = {qf.x, qf.y}, r[2] = {rf.x, rf.y},
                 s[2] = {sf.x, sf.y};
    double c[4][2];
    for (int k = 0; k < 2; ++k) {
      c[0][k] = q[k];
      c[1][k] = -p[k] / 3.0 - q[k] / 2.0 + r[k] - s[k] / 6.0;
      c[2][k] = (p[k] + r[k]) / 2.0 - q[k];
      c[3][k] = (s[k] - p[k]) / 6.0 + (q[k] - r[k]) / 2.0;
    }
    double ph = (dlo[t / M] + (double)m * h) * fcc;
    ph -= rint(ph);
    double sn, cs;
    sincos(2.0 * 3.14159265358979323846 * ph, &sn, &cs);
    const double ar = cs, ai = -sn;
    float o[8];
    for (int k = 0; k < 4; ++k) {
      o[2 * k] = (float)(c[k][0] * ar - c[k][1] * ai);
      o[2 * k + 1] = (float)(c[k][0] * ai + c[k][1] * ar);
    }
    lo = make_float4(o[0], o[1], o[2], o[3]);
    hi = make_float4(o[4], o[5], o[6], o[7]);
  }
  coef[2 * t + 0] = lo;
  coef[2 * t + 1] = hi;
}

// One leg |v - p| from |v|^2 - 2 v.p + |p|^2, with c pointing at the row
// constants -2p and c[-1] = |p|^2:
// the FP64 sum cancels by at most (|v|^2 + |p|^2) / |v - p|^2, a few units
// here, so the leg keeps ~1e-15 relative accuracy.
__device__ __forceinline__ double leg64(double vx, double vy, double vz, double vv,
                                        const double *__restrict__ c) {
  return sqrt_nr64(fma(vz, c[2], fma(vy, c[1], fma(vx, c[0], vv + c[-1]))));
}

// Per (row, voxel).  With vidx, lane t of the launch takes vidx[t] and writes
// output t; without, the brick mapping above and output = linear voxel index.
// blockIdx.y is the row segment (rows [y rps, (y+1) rps)); with more than
// one segment the partial sums go to img / mom + y n for a fixed-order
// reduction.
template <int VPT>
__global__ __launch_bounds__(kBlock) void k_bp_prof(
    float2 *__restrict__ img, float2 *__restrict__ mom, const float *__restrict__ xl,
    const float *__restrict__ yl, const float *__restrict__ zl,
    const int32_t *__restrict__ vidx, int64_t n, int nx, int ny, int nz,
    const double *__restrict__ rowc, const float4 *__restrict__ coef, float thf, int M,
    int rows, int rps) {
  float vx[VPT], vy[VPT], vz[VPT];
  int64_t out[VPT];
  bool ok[VPT];
  if (vidx) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int64_t t = (int64_t)blockIdx.x * (kBlock * VPT) + (int64_t)k * kBlock + threadIdx.x;
      ok[k] = t < n;
      out[k] = t;
      const int64_t lin = vidx[ok[k] ? t : n - 1];
      voxel_coords(xl, yl, zl, ny, nz, lin, vx[k], vy[k], vz[k]);
    }
  } else {
    const int nbz = (nz + 4 * VPT - 1) / (4 * VPT), nby = (ny + 7) / 8;
    const int b = blockIdx.x;
    const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int x = bx * 8 + (w & 1) * 4 + (lane >> 4);
    const int y = by * 8 + (w >> 1) * 4 + ((lane >> 2) & 3);
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const int z = bz * 4 * VPT + k * 4 + (lane & 3);
      ok[k] = x < nx && y < ny && z < nz;
      out[k] = ((int64_t)x * ny + y) * nz + z;
      vx[k] = xl[min(x, nx - 1)];
      vy[k] = yl[min(y, ny - 1)];
      vz[k] = zl[min(z, nz - 1)];
    }
  }
  double dx[VPT], dy[VPT], dz[VPT], vv[VPT];
  float Ir[VPT], Ii[VPT], Cr[VPT], Ci[VPT], Mo[VPT][6];
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    dx[k] = vx[k];
    dy[k] = vy[k];
    dz[k] = vz[k];
    vv[k] = fma(dz[k], dz[k], fma(dy[k], dy[k], dx[k] * dx[k]));
    Ir[k] = Ii[k] = Cr[k] = Ci[k] = 0.0f;
#pragma unroll
    for (int q = 0; q < 6; ++q) Mo[k][q] = 0.0f;
  }
  const float umax = (float)(M - 3);
  const int row_lo = blockIdx.y * rps;
  const int row_hi = min(rows, row_lo + rps);

  // Row i's geometry for every voxel of the lane: the interval's two
  // coefficient vectors (fetched) and the fraction.
  auto fetch = [&](int i, const double *__restrict__ c, float (&f)[VPT], float4 (&a)[VPT],
                   float4 (&b)[VPT]) {
    const bool mono = c[10] != 0.0;
    const float4 *__restrict__ crow = coef + (size_t)i * M * 2;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      // the legs' sum; c[9] doubles it for a monostatic row
      double ds = leg64(dx[k], dy[k], dz[k], vv[k], c + 2);
      if (!mono) ds += leg64(dx[k], dy[k], dz[k], vv[k], c + 6);
      const double u = fma(ds, c[9], c[0]);
      // interval from the clamped FP32 u (NaN-safe), fraction from the FP64
      // u; where FP32 rounding moved m across an integer the fraction lies
      // just outside [0, 1), still inside the cubic's nodes [-1, 2]
      const int m = (int)fminf(fmaxf((float)u, 1.0f), umax);
      f[k] = __builtin_amdgcn_fmed3f((float)(u - (double)m), -1.0f, 2.0f);
      a[k] = crow[2 * m];
      b[k] = crow[2 * m + 1];
    }
  };
  // Row i's contribution from its fetched coefficients: the cubic, the
  // carrier's remainder past the interval start (thf f turns), the block
  // sum and the vote moment.
  auto finish = [&](const double *__restrict__ c, const float (&f)[VPT], const float4 (&a)[VPT],
                    const float4 (&b)[VPT], float (&Sr)[VPT], float (&Si)[VPT]) {
    const float2 va01 = __builtin_bit_cast(float2, c[11]);
    const float vax = va01.x, vay = va01.y, vaz = __builtin_bit_cast(float2, c[12]).x;
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const float qr = fmaf(fmaf(fmaf(b[k].z, f[k], b[k].x), f[k], a[k].z), f[k], a[k].x);
      const float qi = fmaf(fmaf(fmaf(b[k].w, f[k], b[k].y), f[k], a[k].w), f[k], a[k].y);
      const float ph = thf * f[k];
      const float ar = __builtin_amdgcn_cosf(ph), ai = -__builtin_amdgcn_sinf(ph);
      const float cr = fmaf(qr, ar, -(qi * ai));
      const float ci = fmaf(qr, ai, qi * ar);
      Sr[k] += cr;
      Si[k] += ci;
      const float wx = vax - vx[k], wy = vay - vy[k], wz = vaz - vz[k];
      Mo[k][0] = fmaf(wx, cr, Mo[k][0]);
      Mo[k][1] = fmaf(wx, ci, Mo[k][1]);
      Mo[k][2] = fmaf(wy, cr, Mo[k][2]);
      Mo[k][3] = fmaf(wy, ci, Mo[k][3]);
      Mo[k][4] = fmaf(wz, cr, Mo[k][4]);
      Mo[k][5] = fmaf(wz, ci, Mo[k][5]);
    }
  };
  // rows are added in plain FP32 in blocks of kRowBlock, each block's sum
  // then once into the Kahan-compensated total
  auto flush = [&](float (&Sr)[VPT], float (&Si)[VPT]) {
#pragma unroll
    for (int k = 0; k < VPT; ++k) {
      const float yr = Sr[k] - Cr[k], yi = Si[k] - Ci[k];
      const float tr = Ir[k] + yr, ti = Ii[k] + yi;
      Cr[k] = (tr - Ir[k]) - yr;
      Ci[k] = (ti - Ii[k]) - yi;
      Ir[k] = tr;
      Ii[k] = ti;
      Sr[k] = Si[k] = 0.0f;
    }
  };
  float Sr[VPT], Si[VPT];
#pragma unroll
  for (int k = 0; k < VPT; ++k) Sr[k] = Si[k] = 0.0f;
  for (int i = row_lo; i < row_hi; ++i) {
    // the row's constants, all in SGPRs before any is used: one scalar-load
    // wait per row instead of one per use (the compiler otherwise sinks
    // some loads past the first leg and waits twice)
    double rc[13];
    const double *__restrict__ cg = rowc + (size_t)kRowC * i;
#pragma unroll
    for (int k = 0; k < 13; ++k) rc[k] = cg[k];
    asm volatile(""
                 : "+s"(rc[0]), "+s"(rc[1]), "+s"(rc[2]), "+s"(rc[3]), "+s"(rc[4]),
                   "+s"(rc[5]), "+s"(rc[6]), "+s"(rc[7]), "+s"(rc[8]), "+s"(rc[9]),
                   "+s"(rc[10]), "+s"(rc[11]), "+s"(rc[12]));
    float f[VPT];
    float4 a[VPT], b[VPT];
    fetch(i, rc, f, a, b);
    finish(rc, f, a, b, Sr, Si);
    if (((i - row_lo) % kRowBlock) == kRowBlock - 1) flush(Sr, Si);
  }
  flush(Sr, Si);
  const size_t seg = blockIdx.y;
  float2 *__restrict__ im = img + seg * n;
  float2 *__restrict__ mo = mom ? mom + seg * n * 3 : nullptr;
#pragma unroll
  for (int k = 0; k < VPT; ++k) {
    if (!ok[k]) continue;
    const int64_t t = out[k];
    im[t] = make_float2(Ir[k], Ii[k]);
    if (mo) {
      mo[3 * t + 0] = make_float2(Mo[k][0], Mo[k][1]);
      mo[3 * t + 1] = make_float2(Mo[k][2], Mo[k][3]);
      mo[3 * t + 2] = make_float2(Mo[k][4], Mo[k][5]);
    }
  }
}

// --------------------------------------------------------------------------
// Backward of the range-profile forward, for a real loss with g = dL/dI:
//   dL/ds_ij = scale sum_v g_v exp(+j k_j d_iv),
//   dL/dp_i  = scale sum_v Im(conj(g_v) H_i(d_iv)) grad_p d_iv,
//   H_i(d)   = sum_j k_j s_ij exp(-j k_j d),   k_j = 2 pi f_j / c.
// The voxels are taken in groups (bricks of the grid, or runs of a voxel
// list) of centre c_b and radius r_b.  For row i and group b let d0 be the
// round trip to c_b and delta = d - d0, so |delta| <= rho = 2 max_b r_b, and
// xi = delta / rho in [-1, 1].  With k_j = k_c + x_j, x_j = 2 pi (j - t_c)
// df / c, the slow factor exp(-+j x_j delta) of each phasor is a power
// series in xi with argument y_j xi, |y_j| = |x_j| rho <= a:
//   H_i(d0 + delta) = exp(-j k_c delta) sum_k h_k xi^k,
//   h_k = sum_j k_j s_ij exp(-j k_j d0) (-j)^k w_k(y_j),
//   sum_{v in b} g_v exp(+j k_j d_iv)
//       = exp(+j k_j d0) sum_k eps_k j^k J_k(y_j) mu_k,
//   mu_k = sum_{v in b} g_v exp(+j k_c delta_v) T_k(xi_v).
// where exp(-j y xi) = sum_k (-j)^k w_k(y) xi^k is the degree-11 polynomial
// of the Jacobi-Anger sum below turned into powers of xi (w_k(y) = y^k / k!
// would be the power series, which needs 16 terms for the same accuracy).
// The adjoint uses the Chebyshev form exp(+j y xi) = sum_k eps_k j^k J_k(y)
// T_k(xi) (Jacobi-Anger) directly, its moments taken against T_k(xi_v).  Both
// keep 12 terms: the remainder is below 2 sum_{k >= 12} J_k(2) = 4.2e-9 of
// sum_j |.| for a <= 2.  The host sizes the groups for a <= 2; the factors
// w_k(y_j) and eps_k J_k(y_j) depend on j only and come from a table (scalar
// loads).
// k_bwd_taylor gives each lane one row (64 rows per wave) and walks the
// groups of one chunk: per (row, group) the K coefficients h_k from the
// row's samples, per voxel (wave-uniform, through scalar loads) the FP64
// distance, the carrier exp(+j k_c delta) (FP64 phase), the position term
// and the K moments, and per (row, group) the K moments expanded into the
// row's F sample gradients.  Nothing is shared between lanes: no atomics, no
// cross-lane sums; the per-chunk partials are reduced in a fixed order.
// --------------------------------------------------------------------------
constexpr int kTaylorK = 12;      // polynomial terms of the position term
constexpr int kChebK = 12;        // Chebyshev terms of the adjoint
static_assert(kTaylorK <= kChebK, "the position polynomial is built from the Chebyshev terms");
constexpr double kTaylorA = 2.0;  // bound on |y_j xi| the groups are sized for
constexpr int kTJ = 32;           // samples per FP64 re-anchor of exp(-+j k_j d0)

// Per sample j, y_j = 2 pi (j - t_c) df rho / c: eps_k J_k(y_j) (k < kChebK,
// eps_0 = 1, eps_k = 2) in FP64 from the power series of J_k (|y| <= 2: 20
// terms leave < 1e-20), and the power-basis coefficients of the same sum,
//   w_k = sum_{m >= k, m = k mod 2} (-1)^((m - k) / 2) eps_m J_m(y_j) t_mk,
// T_m(xi) = sum_k t_mk xi^k (k < kTaylorK), so that
// sum_k (-j)^k w_k xi^k = sum_m eps_m (-j)^m J_m(y_j) T_m(xi).
__global__ __launch_bounds__(kBlock) void k_bt_tables(float *__restrict__ tay,
                                                      float *__restrict__ bes, int F,
                                                      double ystep, double tc) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= F) return;
  const double y = ((double)j - tc) * ystep;
  const double h = 0.5 * y, h2 = h * h;
  double ej[kChebK];
  double hk = 1.0, kf = 1.0;  // (y/2)^k, k!
  for (int k = 0; k < kChebK; ++k) {
    double term = hk / kf, sum = 0.0;
    for (int m = 0; m < 20; ++m) {
      sum += term;
      term *= -h2 / ((double)(m + 1) * (double)(m + 1 + k));
    }
    ej[k] = (k ? 2.0 : 1.0) * sum;
    bes[(size_t)j * kChebK + k] = (float)ej[k];
    hk *= h;
    kf *= (double)(k + 1);
  }
  // coefficients of T_{m-1} and T_m: T_1 = xi, then T_{m+1} = 2 xi T_m - T_{m-1}
  double tp[kTaylorK] = {}, tm[kTaylorK] = {}, w[kTaylorK] = {};
  tm[0] = 1.0;
  for (int m = 0; m < kTaylorK; ++m) {
    for (int k = m & 1; k <= m; k += 2) w[k] += (((m - k) >> 1) & 1 ? -ej[m] : ej[m]) * tm[k];
    double tn[kTaylorK];
    for (int k = 0; k < kTaylorK; ++k)
      tn[k] = m == 0 ? (k == 1 ? 1.0 : 0.0) : (k ? 2.0 * tm[k - 1] : 0.0) - tp[k];
    for (int k = 0; k < kTaylorK; ++k) {
      tp[k] = tm[k];
      tm[k] = tn[k];
    }
  }
  for (int k = 0; k < kTaylorK; ++k) tay[(size_t)j * kTaylorK + k] = (float)w[k];
}

// Brick of each listed voxel (the sort key) and its list position.
__global__ __launch_bounds__(kBlock) void k_bt_keys(uint32_t *__restrict__ keys,
                                                    int32_t *__restrict__ vals,
                                                    const int32_t *__restrict__ vidx, int64_t n,
                                                    int ny, int nz, int ex, int ey, int ez) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const int64_t lin = vidx[t];
  const int z = (int)(lin % nz), y = (int)((lin / nz) % ny), x = (int)(lin / ((int64_t)ny * nz));
  const int nbz = (nz + ez - 1) / ez, nby = (ny + ey - 1) / ey;
  keys[t] = (uint32_t)(((int64_t)(x / ex) * nby + y / ey) * nbz + z / ez);
  vals[t] = (int32_t)t;
}

__global__ void k_bt_set(int32_t *__restrict__ p, int32_t v) { *p = v; }

// Voxel records in group order.  List mode (vidx): record t is list entry
// perm[t] (the list sorted by brick), total = n.  Grid mode: bricks of
// ex x ey x ez voxels, positions outside the grid padded with weight 0.
__global__ __launch_bounds__(kBlock) void k_bt_pack(
    double4 *__restrict__ recd, float4 *__restrict__ recf, float2 *__restrict__ recg,
    const float2 *__restrict__ g, const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    const int32_t *__restrict__ perm, int64_t n, int nx, int ny, int nz, int ex, int ey,
    int ez, int64_t G, int64_t total) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= total) return;
  float vx, vy, vz;
  float2 gv = make_float2(0.0f, 0.0f);
  if (vidx) {
    const int64_t src = perm[t];
    voxel_coords(xl, yl, zl, ny, nz, vidx[src], vx, vy, vz);
    gv = g[src];
  } else {
    // group b = (bx, by, bz), bz fastest; offset o = (ox, oy, oz), oz fastest
    const int64_t b = t / G, o = t % G;
    const int nbz = (nz + ez - 1) / ez, nby = (ny + ey - 1) / ey;
    const int bz = (int)(b % nbz), by = (int)((b / nbz) % nby), bx = (int)(b / ((int64_t)nbz * nby));
    const int oz = (int)(o % ez), oy = (int)((o / ez) % ey), ox = (int)(o / ((int64_t)ez * ey));
    const int x = bx * ex + ox, y = by * ey + oy, z = bz * ez + oz;
    const bool real = x < nx && y < ny && z < nz;
    vx = xl[min(x, nx - 1)];
    vy = yl[min(y, ny - 1)];
    vz = zl[min(z, nz - 1)];
    if (real) gv = g[((int64_t)x * ny + y) * nz + z];
  }
  const double dx = vx, dy = vy, dz = vz;
  recd[t] = make_double4(dx, dy, dz, fma(dz, dz, fma(dy, dy, dx * dx)));
  recf[t] = make_float4(vx, vy, vz, 0.0f);
  recg[t] = gv;
}

// One workgroup per group (G records each, or [goff[b], goff[b+1])): the
// midpoint of its box and its radius about it (FP64); the largest radius also
// goes to *rmax (as the bits of a positive double, ordered like the values).
__global__ __launch_bounds__(kBlock) void k_bt_groups(double4 *__restrict__ groupc,
                                                      unsigned long long *__restrict__ rmax,
                                                      const double4 *__restrict__ recd,
                                                      int64_t G,
                                                      const int32_t *__restrict__ goff) {
  __shared__ double red[6][kBlock / 64];
  const int64_t b = blockIdx.x;
  const int64_t b0 = goff ? (int64_t)goff[b] : b * G;
  const double4 *__restrict__ r = recd + b0;
  if (goff) G = (int64_t)goff[b + 1] - b0;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t t = threadIdx.x; t < G; t += kBlock) {
    const double4 v = r[t];
    lo[0] = fmin(lo[0], v.x), hi[0] = fmax(hi[0], v.x);
    lo[1] = fmin(lo[1], v.y), hi[1] = fmax(hi[1], v.y);
    lo[2] = fmin(lo[2], v.z), hi[2] = fmax(hi[2], v.z);
  }
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  for (int k = 0; k < 3; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, 64));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, 64));
    }
    if (lane == 0) {
      red[k][w] = lo[k];
      red[3 + k][w] = hi[k];
    }
  }
  __syncthreads();
  double c[3];
  for (int k = 0; k < 3; ++k) {
    double l = red[k][0], h = red[3 + k][0];
    for (int q = 1; q < kBlock / 64; ++q) l = fmin(l, red[k][q]), h = fmax(h, red[3 + k][q]);
    c[k] = 0.5 * (l + h);
  }
  __syncthreads();
  double rr = 0.0;
  for (int64_t t = threadIdx.x; t < G; t += kBlock) {
    const double4 v = r[t];
    const double ex = v.x - c[0], ey = v.y - c[1], ez = v.z - c[2];
    rr = fmax(rr, sqrt(fma(ez, ez, fma(ey, ey, ex * ex))));
  }
  for (int o = 32; o > 0; o >>= 1) rr = fmax(rr, __shfl_xor(rr, o, 64));
  if (lane == 0) red[0][w] = rr;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kBlock / 64; ++q) rr = fmax(rr, red[0][q]);
    groupc[b] = make_double4(c[0], c[1], c[2], rr);
    atomicMax(rmax, (unsigned long long)__double_as_longlong(rr));
  }
}

// k_j s_ij, transposed to [j][i] so that a wave's 64 rows read one sample
// with one coalesced load.
__global__ __launch_bounds__(kBlock) void k_bt_meas(float2 *__restrict__ mt,
                                                    const float2 *__restrict__ meas, int A,
                                                    int F, float k0, float dk) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)A * F) return;
  const int j = (int)(t / A), i = (int)(t % A);
  const float2 s = meas[(size_t)i * F + j];
  const float kj = fmaf((float)j, dk, k0);
  mt[t] = make_float2(kj * s.x, kj * s.y);
}

// Lane = row.  BI: some row has a nonzero RX offset (then both legs are
// evaluated for every row; a monostatic row's second leg is its first).
// Monostatic: five waves per SIMD (94 VGPRs, no spills) instead of the four
// that 112 VGPRs allow, 5 % faster at config 4.  Bistatic (both legs) keeps
// four: at five it spills.
template <int K, int KC, bool BI>
__global__ __launch_bounds__(kBlock, BI ? 4 : 5) void k_bwd_taylor(
    float2 *__restrict__ st, float *__restrict__ pp, const double4 *__restrict__ recd,
    const float4 *__restrict__ recf, const float2 *__restrict__ recg,
    const double4 *__restrict__ groupc, int64_t G, const int32_t *__restrict__ goff,
    int ngroups, int gpc, const float *__restrict__ ant, const float *__restrict__ rxo,
    const float2 *__restrict__ mt, const float *__restrict__ tay,
    const float *__restrict__ bes, double f0c, double dfc, double tc, double rho, int A,
    int R, int F, int is_ti) {
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  // a wave whose 64 rows all lie past A has nothing to do (no barriers here)
  if ((blockIdx.x * 4 + wv) * 64 >= A) return;
  const int i_raw = (blockIdx.x * 4 + wv) * 64 + (int)(threadIdx.x % 64);
  const bool live = i_raw < A;
  const int i = live ? i_raw : A - 1;
  const int c = blockIdx.y;
  const int g_lo = c * gpc, g_hi = min(ngroups, g_lo + gpc);

  const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
  const int rx = i % R;
  const float qx = px + rxo[3 * rx + 0], qy = py + rxo[3 * rx + 1], qz = pz + rxo[3 * rx + 2];
  const double pdx = px, pdy = py, pdz = pz, qdx = qx, qdy = qy, qdz = qz;
  const double p2 = fma(pdz, pdz, fma(pdy, pdy, pdx * pdx));
  const double q2 = fma(qdz, qdz, fma(qdy, qdy, qdx * qdx));
  const double ti = is_ti ? 0.15 : 0.0;
  const double fcc = fma(tc, dfc, f0c);            // f_c / c, turns per metre
  const double inv_rho = 1.0 / rho;
  // delta = two * (leg sum as evaluated) + ti - d0; the monostatic sum is one leg
  const double two = BI ? 1.0 : 2.0;
  const double xs = two * inv_rho, ps = two * fcc;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;

  for (int b = g_lo; b < g_hi; ++b) {
    const double4 cen = groupc[b];
    // d0: the row's round trip to the group centre
    const double e1x = cen.x - pdx, e1y = cen.y - pdy, e1z = cen.z - pdz;
    double d0 = sqrt(fma(e1z, e1z, fma(e1y, e1y, e1x * e1x)));
    if (BI) {
      const double e2x = cen.x - qdx, e2y = cen.y - qdy, e2z = cen.z - qdz;
      d0 += sqrt(fma(e2z, e2z, fma(e2y, e2y, e2x * e2x)));
    } else {
      d0 += d0;
    }
    d0 += ti;
    const double xo = (ti - d0) * inv_rho, po = (ti - d0) * fcc;
    float sr, si;  // exp(-j k_j d0) steps by exp(-j 2 pi df d0 / c)
    phasor_quarter(4.0 * d0 * dfc, sr, si);

    // h_k = sum_j k_j s_ij exp(-j k_j d0) (-j)^k w_k(y_j)
    float hr[K], hi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) hr[k] = hi[k] = 0.0f;
    for (int j0 = 0; j0 < F; j0 += kTJ) {
      float er, ei;
      phasor_hw(d0 * fma((double)j0, dfc, f0c), er, ei);
      const int j1 = min(F, j0 + kTJ);
      for (int j = j0; j < j1; ++j) {
        const float2 a = mt[(size_t)j * A + i];
        const float tr = fmaf(a.x, er, -(a.y * ei)), tq = fmaf(a.x, ei, a.y * er);
        const float *__restrict__ tw = tay + (size_t)j * K;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          // (-j)^k t: k mod 4 = 0: t, 1: (tq, -tr), 2: -t, 3: (-tq, tr)
          const float ur = (k & 1) ? tq : tr, uq = (k & 1) ? tr : tq;
          const float sgr = ((k & 3) == 2 || (k & 3) == 3) ? -1.0f : 1.0f;
          const float sgq = ((k & 3) == 1 || (k & 3) == 2) ? -1.0f : 1.0f;
          hr[k] = fmaf(sgr * tw[k], ur, hr[k]);
          hi[k] = fmaf(sgq * tw[k], uq, hi[k]);
        }
        const float nr = fmaf(er, sr, -(ei * si)), ni = fmaf(er, si, ei * sr);
        er = nr;
        ei = ni;
      }
    }

    float mr[KC], mi[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) mr[k] = mi[k] = 0.0f;
    // one voxel's work: FP64 round trip, carrier, position term, moments
    auto voxel = [&](const double4 &rd, const float4 &rf, float gr, float gi) {
      const double s1 =
          fma(rd.z, -2.0 * pdz, fma(rd.y, -2.0 * pdy, fma(rd.x, -2.0 * pdx, rd.w + p2)));
      const float y1 = __builtin_amdgcn_rsqf((float)s1);
      const double y1d = y1, a1 = s1 * y1d;
      double d = fma(0.5 * y1d, fma(-a1, a1, s1), a1);
      float y2 = y1;
      if (BI) {
        const double s2 =
            fma(rd.z, -2.0 * qdz, fma(rd.y, -2.0 * qdy, fma(rd.x, -2.0 * qdx, rd.w + q2)));
        y2 = __builtin_amdgcn_rsqf((float)s2);
        const double y2d = y2, a2 = s2 * y2d;
        d += fma(0.5 * y2d, fma(-a2, a2, s2), a2);
      }
      const float xi = (float)fma(d, xs, xo);
      const float ph = (float)__builtin_amdgcn_fract(fma(d, ps, po));
      const float cs = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
      // c = g exp(+j k_c delta)
      const float cr = fmaf(gr, cs, -(gi * sn)), ci = fmaf(gr, sn, gi * cs);
      // position term: Im(conj(g) H) = Im(conj(c) sum_k h_k xi^k)
      float Pr = hr[K - 1], Pi = hi[K - 1];
#pragma unroll
      for (int k = K - 2; k >= 0; --k) {
        Pr = fmaf(Pr, xi, hr[k]);
        Pi = fmaf(Pi, xi, hi[k]);
      }
      const float tt = fmaf(cr, Pi, -(ci * Pr));
      if (BI) {
        const float w1 = tt * y1, w2 = tt * y2;
        gx = fmaf(w1, px - rf.x, fmaf(w2, qx - rf.x, gx));
        gy = fmaf(w1, py - rf.y, fmaf(w2, qy - rf.y, gy));
        gz = fmaf(w1, pz - rf.z, fmaf(w2, qz - rf.z, gz));
      } else {
        const float w1 = tt * y1;  // the factor 2 of the two legs is applied at the end
        gx = fmaf(w1, px - rf.x, gx);
        gy = fmaf(w1, py - rf.y, gy);
        gz = fmaf(w1, pz - rf.z, gz);
      }
      // adjoint moments mu_k += c T_k(xi)
      float t0 = 1.0f, t1 = xi;
      mr[0] += cr;
      mi[0] += ci;
#pragma unroll
      for (int k = 1; k < KC; ++k) {
        mr[k] = fmaf(cr, t1, mr[k]);
        mi[k] = fmaf(ci, t1, mi[k]);
        const float t2 = fmaf(2.0f * xi, t1, -t0);
        t0 = t1;
        t1 = t2;
      }
    };
    // Voxels one at a time: measured on MI355X, two per iteration (more
    // independent chains, 156 VGPRs), even/odd splits of the Horner and
    // Chebyshev chains (half their depth) and prefetching the next record
    // gained nothing; pipelining the records' scalar loads one voxel deep is undone
    // by the scheduler, which sinks the loads back to their use.  The empty
    // asm takes the whole record as in/out SGPR operands, so all of it is
    // loaded before any is used: one scalar-load wait per voxel, not two.
    const int64_t v0 = goff ? (int64_t)goff[b] : (int64_t)b * G;
    const int64_t v1 = goff ? (int64_t)goff[b + 1] : v0 + G;
    for (int64_t v = v0; v < v1; ++v) {
      const double4 rd = recd[v];
      const float4 rf = recf[v];
      const float2 gv = recg[v];
      double cx = rd.x, cy = rd.y, cz = rd.z, cw = rd.w;
      float fx = rf.x, fy = rf.y, fz = rf.z, gr = gv.x, gi = gv.y;
      asm volatile(""
                   : "+s"(cx), "+s"(cy), "+s"(cz), "+s"(cw), "+s"(fx), "+s"(fy), "+s"(fz),
                     "+s"(gr), "+s"(gi));
      voxel(make_double4(cx, cy, cz, cw), make_float4(fx, fy, fz, 0.0f), gr, gi);
    }

    // S_ij += exp(+j k_j d0) sum_k eps_k j^k J_k(y_j) mu_k  (conjugate steps)
    float2 *__restrict__ srow = st + (size_t)c * F * A + i;
    for (int j0 = 0; j0 < F; j0 += kTJ) {
      float er, ei;
      phasor_hw(d0 * fma((double)j0, dfc, f0c), er, ei);
      ei = -ei;
      const int j1 = min(F, j0 + kTJ);
      for (int j = j0; j < j1; ++j) {
        // sum_k eps_k J_k(y_j) i^k mu_k; i^k mu: k mod 4 = 0: mu, 1: (-mi, mr),
        // 2: -mu, 3: (mi, -mr)
        const float *__restrict__ bw = bes + (size_t)j * KC;
        float tr = 0.0f, tq = 0.0f;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          const float ur = (k & 1) ? mi[k] : mr[k], uq = (k & 1) ? mr[k] : mi[k];
          const float sgr = ((k & 3) == 1 || (k & 3) == 2) ? -1.0f : 1.0f;
          const float sgq = ((k & 3) == 2 || (k & 3) == 3) ? -1.0f : 1.0f;
          tr = fmaf(sgr * bw[k], ur, tr);
          tq = fmaf(sgq * bw[k], uq, tq);
        }
        const float vr = fmaf(er, tr, -(ei * tq)), vq = fmaf(er, tq, ei * tr);
        if (live) {
          float2 *__restrict__ o = srow + (size_t)j * A;
          if (b == g_lo) {
            *o = make_float2(vr, vq);
          } else {
            const float2 old = *o;
            *o = make_float2(old.x + vr, old.y + vq);
          }
        }
        const float nr = fmaf(er, sr, ei * si), ni = fmaf(ei, sr, -(er * si));
        er = nr;
        ei = ni;
      }
    }
  }
  if (live) {
    const float two_legs = BI ? 1.0f : 2.0f;  // exact: a power of two
    float *__restrict__ o = pp + ((size_t)c * A + i) * 3;
    o[0] = two_legs * gx;
    o[1] = two_legs * gy;
    o[2] = two_legs * gz;
  }
}

// grad_meas[i][j] = scale sum_c st[c][j][i] (FP64, fixed order).
// grad[i][j] = scale * sum_q st[q][j][i], q in order (FP64).  A block takes
// a tile of 64 rows x 8 samples: the chunk partials are read along i
// (coalesced; each thread keeps two independent sums in flight), the tile is
// transposed through LDS and written along j.
constexpr int kRedI = 64, kRedJ = 8;
__global__ __launch_bounds__(kBlock) void k_bt_reduce_meas(float2 *__restrict__ grad,
                                                           const float2 *__restrict__ st,
                                                           int chunks, int A, int F,
                                                           double scale) {
  constexpr int kRows = kBlock / kRedI;  // 4 sample rows of the tile per pass
  constexpr int kPer = kRedJ / kRows;    // 2 samples per thread
  __shared__ float2 tile[kRedJ][kRedI + 1];
  const int i0 = blockIdx.x * kRedI, j0 = blockIdx.y * kRedJ;
  const int64_t AF = (int64_t)A * F;
  const int ii = threadIdx.x % kRedI, jb = threadIdx.x / kRedI;
  const int i = i0 + ii;
  double a[kPer] = {}, b[kPer] = {};
  bool ok[kPer];
  const float2 *src[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int j = j0 + jb + k * kRows;
    ok[k] = i < A && j < F;
    src[k] = st + (size_t)(ok[k] ? j : 0) * A + (ok[k] ? i : 0);
  }
  // four chunks' loads in flight per sample, added in chunk order
  constexpr int kU = 4;
  int q = 0;
  for (; q + kU <= chunks; q += kU) {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      if (!ok[k]) continue;
      float2 v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = src[k][(size_t)(q + u) * AF];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        a[k] += v[u].x;
        b[k] += v[u].y;
      }
    }
  }
  for (; q < chunks; ++q) {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      if (!ok[k]) continue;
      const float2 v = src[k][(size_t)q * AF];
      a[k] += v.x;
      b[k] += v.y;
    }
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k)
    tile[jb + k * kRows][ii] = make_float2((float)(scale * a[k]), (float)(scale * b[k]));
  __syncthreads();
  for (int e = threadIdx.x; e < kRedI * kRedJ; e += kBlock) {
    const int jj = e % kRedJ, iq = e / kRedJ;
    const int io = i0 + iq, j = j0 + jj;
    if (io < A && j < F) grad[(size_t)io * F + j] = tile[jj][iq];
  }
}

__global__ __launch_bounds__(kBlock) void k_bt_reduce_ant(float *__restrict__ grad,
                                                          const float *__restrict__ pp,
                                                          int chunks, int A, double scale) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= 3 * A) return;
  double s = 0.0;
  // eight chunks' loads in flight, added in chunk order
  constexpr int kU = 8;
  int q = 0;
  for (; q + kU <= chunks; q += kU) {
    float v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = pp[(size_t)(q + u) * A * 3 + t];
#pragma unroll
    for (int u = 0; u < kU; ++u) s += v[u];
  }
  for (; q < chunks; ++q) s += pp[(size_t)q * A * 3 + t];
  grad[t] = (float)(scale * s);
}

// Group shape: bricks whose edges (in voxels) are balanced over each axis
// and at most emax.  The grid is packed brick by brick; a voxel list is
// sorted by brick and each non-empty brick is one group.
struct BtShape {
  int ex, ey, ez;
  int64_t G, ngroups, total;
};

BtShape bt_shape(int nx, int ny, int nz, int emax) {
  BtShape s;
  auto edge = [&](int nax) {
    const int nb = (nax + emax - 1) / emax;
    return (nax + nb - 1) / nb;
  };
  s.ex = edge(nx);
  s.ey = edge(ny);
  s.ez = edge(nz);
  s.G = (int64_t)s.ex * s.ey * s.ez;
  s.ngroups = (int64_t)((nx + s.ex - 1) / s.ex) * ((ny + s.ey - 1) / s.ey) *
              ((nz + s.ez - 1) / s.ez);
  s.total = s.G * s.ngroups;
  return s;
}

// Temporary storage of the list sort: radix sort by brick, run-length
// encoding of the sorted bricks, exclusive scan of the run lengths.
size_t bt_cub_bytes(int64_t n) {
  size_t a = 0, b = 0, c = 0;
  const int m = (int)n;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint32_t *)nullptr,
                                          (uint32_t *)nullptr, (const int32_t *)nullptr,
                                          (int32_t *)nullptr, m, 0, 32, (hipStream_t)0);
  (void)hipcub::DeviceRunLengthEncode::Encode(nullptr, b, (const uint32_t *)nullptr,
                                             (uint32_t *)nullptr, (int32_t *)nullptr,
                                             (int32_t *)nullptr, m, (hipStream_t)0);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const int32_t *)nullptr,
                                        (int32_t *)nullptr, m, (hipStream_t)0);
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

struct BtWs {
  double4 *recd, *groupc;
  float4 *recf;
  float2 *recg, *mt, *st;
  float *pp, *tay, *bes;
  // the list sort (list mode only)
  uint32_t *k_in, *k_out, *uniq;
  int32_t *v_in, *v_out, *cnt, *goff, *nruns;
  void *cub;
  size_t cub_bytes;
  unsigned long long *rmax;
};

// Most chunks (partial copies of the sample gradients) a launch may use.
int bt_chunks_max(int A, int F) {
  int64_t c = ((int64_t)1 << 30) / ((int64_t)A * F * 8);
  if (c > 512) c = 512;
  return c < 1 ? 1 : (int)c;
}

size_t bt_layout(int A, int F, int64_t total, int64_t ngroups, int64_t nlist, BtWs *w,
                 char *base) {
  const int cm = bt_chunks_max(A, F);
  const size_t cubb = nlist > 0 ? bt_cub_bytes(nlist) : 0;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  char *rd = take(sizeof(double4) * (size_t)total);
  char *rf = take(sizeof(float4) * (size_t)total);
  char *rg = take(sizeof(float2) * (size_t)total);
  char *gc = take(sizeof(double4) * (size_t)ngroups);
  char *m = take(sizeof(float2) * (size_t)A * F);
  char *s = take(sizeof(float2) * (size_t)cm * A * F);
  char *p = take(sizeof(float) * 3 * (size_t)cm * A);
  char *r = take(sizeof(unsigned long long));
  char *ty = take(sizeof(float) * kTaylorK * (size_t)F);
  char *bs = take(sizeof(float) * kChebK * (size_t)F);
  const size_t nl = nlist > 0 ? (size_t)nlist : 0;
  char *ki = take(sizeof(uint32_t) * nl), *ko = take(sizeof(uint32_t) * nl);
  char *vi = take(sizeof(int32_t) * nl), *vo = take(sizeof(int32_t) * nl);
  char *un = take(sizeof(uint32_t) * nl), *ct = take(sizeof(int32_t) * nl);
  char *go = take(sizeof(int32_t) * (nl + 1)), *nr = take(sizeof(int32_t));
  char *cb = take(cubb);
  if (w) {
    w->k_in = (uint32_t *)ki;
    w->k_out = (uint32_t *)ko;
    w->v_in = (int32_t *)vi;
    w->v_out = (int32_t *)vo;
    w->uniq = (uint32_t *)un;
    w->cnt = (int32_t *)ct;
    w->goff = (int32_t *)go;
    w->nruns = (int32_t *)nr;
    w->cub = cb;
    w->cub_bytes = cubb;
    w->tay = (float *)ty;
    w->bes = (float *)bs;
    w->recd = (double4 *)rd;
    w->recf = (float4 *)rf;
    w->recg = (float2 *)rg;
    w->groupc = (double4 *)gc;
    w->mt = (float2 *)m;
    w->st = (float2 *)s;
    w->pp = (float *)p;
    w->rmax = (unsigned long long *)r;
  }
  return off;
}

// Largest brick edge (voxels) whose radius keeps |y_j xi| <= kTaylorA, from
// the axis pitches (the group radii are measured again on the device).
int bt_emax(int F, double df, const double *bb, int nx, int ny, int nz) {
  const double band = 0.5 * (double)(F - 1) * fabs(df);  // t_c |df|
  if (band <= 0.0) return 64;
  // rho = 2 r, a = 2 pi band rho / c  ->  r <= kTaylorA c / (4 pi band)
  const double rb = kTaylorA * MMN_C0 / (4.0 * 3.141592653589793 * band);
  const double px = nx > 1 ? (bb[3] - bb[0]) / (nx - 1) : 0.0;
  const double py = ny > 1 ? (bb[4] - bb[1]) / (ny - 1) : 0.0;
  const double pz = nz > 1 ? (bb[5] - bb[2]) / (nz - 1) : 0.0;
  const double pm = fmax(px, fmax(py, pz));
  if (pm <= 0.0) return 64;
  int e = (int)floor(2.0 * 0.95 * rb / (sqrt(3.0) * pm)) + 1;
  if (e > 64) e = 64;
  return e < 1 ? 1 : e;
}

int env_int(const char *name, int dflt) {
  const char *s = getenv(name);
  return (s && *s) ? atoi(s) : dflt;
}

}  // namespace

extern "C" size_t mmn_bp_profile_workspace_size(int64_t n_vox, int num_rows, int samples,
                                                double f0, double df, const double *bbox,
                                                int with_moment) {
  ProfGeom g;
  if (n_vox <= 0 || num_rows <= 0 || samples <= 0 || !bbox ||
      !prof_geom(samples, f0, df, bbox, &g))
    return 0;
  return prof_layout(num_rows, g.M, n_vox, with_moment != 0, nullptr, nullptr);
}

extern "C" int mmn_bp_profile(float *img, float *moment, const float *x_locs,
                              const float *y_locs, const float *z_locs,
                              const int32_t *vox_idx, int64_t n_vox, const float *ant,
                              const float *meas, const float *rx_off, double f0, double df,
                              int nx, int ny, int nz, int num_rows, int num_rx, int samples,
                              int is_ti_radar, const double *bbox, void *workspace,
                              size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(img && x_locs && y_locs && z_locs && ant && meas && rx_off && bbox,
              MMN_ERR_NULL, "mmn_bp_profile: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 && samples > 0,
              MMN_ERR_ARG, "mmn_bp_profile: sizes must be positive");
  ProfGeom g;
  MMN_REQUIRE(prof_geom(samples, f0, df, bbox, &g), MMN_ERR_ARG,
              "mmn_bp_profile: f0, df and the bounding box must be finite (box min <= "
              "max), and the box's round-trip span at most %d table points", kMaxM);
  int64_t n = (int64_t)nx * ny * nz;
  if (vox_idx) {
    MMN_REQUIRE(n_vox >= 0, MMN_ERR_ARG, "mmn_bp_profile: n_vox < 0");
    n = n_vox;
  }
  if (n == 0) return 0;
  const int A = num_rows, F = samples, M = g.M;
  const bool with_mom = moment != nullptr;
  const size_t need = prof_layout(A, M, n, with_mom, nullptr, nullptr);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_bp_profile: workspace of %zu bytes needed", need);
  ProfWs w;
  prof_layout(A, M, n, with_mom, &w, (char *)workspace);
  hipStream_t st = (hipStream_t)stream;
  Box box;
  for (int k = 0; k < 6; ++k) box.b[k] = bbox[k];
  const double inv_c = 1.0 / MMN_C0;
  const double tc = 0.5 * (double)(F - 1);
  const double fcc = (f0 + tc * df) * inv_c;
  hipLaunchKernelGGL(k_prof_rows, dim3(mmn_blocks(A, kBlock)), dim3(kBlock), 0, st, w.dlo,
                     w.rowc, ant, rx_off, box, A, num_rx, is_ti_radar, g.h);
  hipLaunchKernelGGL(k_prof_table, dim3(mmn_blocks(M, kBlock), A), dim3(kBlock), 0, st, w.tab,
                     (const float2 *)meas, w.dlo, g.h, df * inv_c, tc, M, F, 0, 0.0f, 0.0f);
  hipLaunchKernelGGL(k_prof_coef, dim3(mmn_blocks((int64_t)A * M, kBlock)), dim3(kBlock), 0,
                     st, w.coef, w.tab, w.dlo, g.h, fcc, M, (int64_t)A * M);

  // Two voxels per lane (an 8x8x8 brick per workgroup) unless MMN_PROF_VPT
  // says 1.  Measured on MI355X (config 2 / config 5): 8.5 / 124 ms at two,
  // 9.4 / 145 ms at one, 9.0 / 135 ms at four; fetching row i + 1's
  // coefficients while row i finishes (software pipelining) cost 7-25 %.
  const int vpt = env_int("MMN_PROF_VPT", 2) == 1 ? 1 : 2;
  auto kern = vpt == 1 ? k_bp_prof<1> : k_bp_prof<2>;
  int64_t nblk;
  if (vox_idx)
    nblk = mmn_blocks(n, kBlock * vpt);
  else
    nblk = (int64_t)((nx + 7) / 8) * ((ny + 7) / 8) * ((nz + 4 * vpt - 1) / (4 * vpt));
  // row segments: enough workgroups for eight rounds of resident slots
  int per_cu = 0, dev = 0, cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 2;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipGetLastError();
  int rs = 1;
  const int forced = env_int("MMN_PROF_RS", 0);
  if (forced > 0) {
    rs = forced;
  } else {
    while (nblk * rs < 8 * (int64_t)per_cu * cus && rs < kMaxSeg) ++rs;
  }
  if (rs > seg_cap(A)) rs = seg_cap(A);
  const int rps = (A + rs - 1) / rs;
  const int nseg = (A + rps - 1) / rps;
  float2 *im = (float2 *)img, *mo = (float2 *)moment;
  if (nseg > 1) {
    im = w.part;
    mo = with_mom ? w.part + (size_t)nseg * n : nullptr;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)nblk, nseg), dim3(kBlock), 0, st, im, mo, x_locs,
                     y_locs, z_locs, vox_idx, n, nx, ny, nz, w.rowc,
                     (const float4 *)w.coef, (float)(fcc * g.h), M, A, rps);
  if (nseg > 1) mmn_rs_reduce(st, (float2 *)img, (float2 *)moment, im, mo, n, nseg);
  return mmn_check_launch("mmn_bp_profile");
}

extern "C" size_t mmn_bp_profile_backward_workspace_size(int64_t n_vox, int num_rows,
                                                         int samples, double f0, double df,
                                                         const double *bbox) {
  (void)f0;
  (void)df;
  if (n_vox <= 0 || n_vox >= ((int64_t)1 << 31) || num_rows <= 0 || samples <= 0 || !bbox)
    return 0;
  // records: a grid's bricks pad each axis by less than a factor of two
  // (< 8 n); a list needs n and the sort buffers; groups: at most n
  const int64_t total = n_vox * 8, ngroups = n_vox;
  return bt_layout(num_rows, samples, total, ngroups, n_vox, nullptr, nullptr);
}

extern "C" int mmn_bp_profile_backward(float *grad_meas, float *grad_ant, const float *g,
                                       double scale, const float *x_locs, const float *y_locs,
                                       const float *z_locs, const int32_t *vox_idx,
                                       int64_t n_vox, const float *ant, const float *meas,
                                       const float *rx_off, double f0, double df, int nx,
                                       int ny, int nz, int num_rows, int num_rx, int samples,
                                       int is_ti_radar, const double *bbox, void *workspace,
                                       size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(g && x_locs && y_locs && z_locs && ant && meas && rx_off && bbox, MMN_ERR_NULL,
              "mmn_bp_profile_backward: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 && samples > 0,
              MMN_ERR_ARG, "mmn_bp_profile_backward: sizes must be positive");
  MMN_REQUIRE(isfinite(f0) && isfinite(df) && isfinite(bbox[0] + bbox[1] + bbox[2] + bbox[3] +
                                                       bbox[4] + bbox[5]),
              MMN_ERR_ARG, "mmn_bp_profile_backward: f0, df and bbox must be finite");
  int64_t n = (int64_t)nx * ny * nz;
  if (vox_idx) {
    MMN_REQUIRE(n_vox >= 0, MMN_ERR_ARG, "mmn_bp_profile_backward: n_vox < 0");
    n = n_vox;
  }
  const int A = num_rows, F = samples;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0 || (!grad_meas && !grad_ant)) {
    int rc = 0;
    if (grad_meas)
      rc = mmn_zero_async(grad_meas, sizeof(float2) * (size_t)A * F, st,
                          "mmn_bp_profile_backward");
    if (rc == 0 && grad_ant)
      rc = mmn_zero_async(grad_ant, sizeof(float) * 3 * (size_t)A, st,
                          "mmn_bp_profile_backward");
    return rc;
  }
  const size_t need = mmn_bp_profile_backward_workspace_size(n, A, F, f0, df, bbox);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_bp_profile_backward: workspace of %zu bytes needed", need);
  const double inv_c = 1.0 / MMN_C0;
  const double two_pi = 6.283185307179586476925286766559;
  const double tc = 0.5 * (double)(F - 1);
  const double band = tc * fabs(df);  // t_c |df|
  int emax = bt_emax(F, df, bbox, nx, ny, nz);
  // Pack the voxels in group order and measure the groups; halve the brick
  // edge until the largest radius satisfies the series bound (the pitch
  // estimate above assumes evenly spaced axes).
  BtShape sh;
  BtWs w;
  double rho = 0.0;
  const int32_t *goff = nullptr;
  for (;;) {
    sh = bt_shape(nx, ny, nz, emax);
    hipError_t e = hipSuccess;
    if (vox_idx) {
      // sort the list by brick; each run of one brick is a group
      bt_layout(A, F, n, n, n, &w, (char *)workspace);
      hipLaunchKernelGGL(k_bt_keys, dim3(mmn_blocks(n, kBlock)), dim3(kBlock), 0, st, w.k_in,
                         w.v_in, vox_idx, n, ny, nz, sh.ex, sh.ey, sh.ez);
      int bits = 1;
      while (bits < 32 && ((int64_t)1 << bits) < sh.ngroups) ++bits;
      size_t tb = w.cub_bytes;
      e = hipcub::DeviceRadixSort::SortPairs(w.cub, tb, w.k_in, w.k_out, w.v_in, w.v_out, (int)n,
                                             0, bits, st);
      tb = w.cub_bytes;
      if (e == hipSuccess)
        e = hipcub::DeviceRunLengthEncode::Encode(w.cub, tb, w.k_out, w.uniq, w.cnt, w.nruns,
                                                  (int)n, st);
      int32_t runs = 0;
      if (e == hipSuccess) e = hipMemcpyAsync(&runs, w.nruns, sizeof(runs), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      tb = w.cub_bytes;
      if (e == hipSuccess)
        e = hipcub::DeviceScan::ExclusiveSum(w.cub, tb, w.cnt, w.goff, runs, st);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_bt_set, dim3(1), dim3(1), 0, st, w.goff + runs, (int32_t)n);
      sh.ngroups = runs;
      sh.total = n;
      goff = w.goff;
    } else {
      bt_layout(A, F, sh.total, sh.ngroups, 0, &w, (char *)workspace);
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_bt_pack, dim3(mmn_blocks(sh.total, kBlock)), dim3(kBlock), 0, st,
                         w.recd, w.recf, w.recg, (const float2 *)g, x_locs, y_locs, z_locs,
                         vox_idx, w.v_out, n, nx, ny, nz, sh.ex, sh.ey, sh.ez, sh.G, sh.total);
      const int zr =
          mmn_zero_async(w.rmax, sizeof(unsigned long long), st, "mmn_bp_profile_backward");
      if (zr != 0) return zr;
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_bt_groups, dim3((unsigned)sh.ngroups), dim3(kBlock), 0, st,
                         w.groupc, w.rmax, w.recd, sh.G, goff);
      unsigned long long bits = 0;
      e = hipMemcpyAsync(&bits, w.rmax, sizeof(bits), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      double r;
      memcpy(&r, &bits, sizeof(r));
      rho = 2.0 * r;
    }
    if (e != hipSuccess) {
      mmn_set_error("mmn_bp_profile_backward: %s", hipGetErrorString(e));
      return (int)e;
    }
    const double a = two_pi * band * rho * inv_c;
    if (a <= kTaylorA * (1.0 + 1e-9) || emax == 1) break;
    emax = emax / 2;
  }
  MMN_REQUIRE(two_pi * band * rho * inv_c <= kTaylorA * (1.0 + 1e-9), MMN_ERR_ARG,
              "mmn_bp_profile_backward: voxel spacing too coarse for the band (a single "
              "voxel group exceeds the series bound)");
  // single-voxel groups (xi = 0 throughout): any rho within the bound keeps
  // |y_j| <= a and the tables' series accurate
  if (!(rho > 0.0)) rho = band > 0.0 ? kTaylorA * MMN_C0 / (two_pi * band) : 1.0;
  hipLaunchKernelGGL(k_bt_meas, dim3(mmn_blocks((int64_t)A * F, kBlock)), dim3(kBlock), 0, st,
                     w.mt, (const float2 *)meas, A, F, (float)(two_pi * f0 * inv_c),
                     (float)(two_pi * df * inv_c));
  // chunks of groups: ~16 blocks (64 waves) per CU, at most the workspace's
  // (1 GiB of chunk partials, 512 chunks).  Short waves keep the tail small:
  // measured on MI355X, 128 chunks at config 2 (A = 3600; the workspace's
  // limit is 145) and 343 (one group each) at config 4 (A = 1600) beat
  // 64 / 115 by 5 % / 10 %; the chunk sums cost ~0.2 ms.
  bool bi = false;
  {
    // any RX offset nonzero -> both legs (host copy of a few floats)
    float ro[3 * 64];
    const int nr = num_rx < 64 ? num_rx : 64;
    hipError_t e = hipMemcpyAsync(ro, rx_off, sizeof(float) * 3 * nr, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_bp_profile_backward: %s", hipGetErrorString(e));
      return (int)e;
    }
    for (int k = 0; k < 3 * nr; ++k) bi |= ro[k] != 0.0f;
    if (num_rx > 64) bi = true;
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipGetLastError();
  const int64_t rowblocks = (A + 255) / 256;
  int64_t ch = (16 * (int64_t)cus + rowblocks - 1) / rowblocks;
  const int forced = env_int("MMN_PROFB_CHUNKS", 0);
  if (forced > 0) ch = forced;
  if (ch > sh.ngroups) ch = sh.ngroups;
  if (ch > bt_chunks_max(A, F)) ch = bt_chunks_max(A, F);
  if (ch < 1) ch = 1;
  const int gpc = (int)((sh.ngroups + ch - 1) / ch);
  ch = (sh.ngroups + gpc - 1) / gpc;
  hipLaunchKernelGGL(k_bt_tables, dim3(mmn_blocks(F, kBlock)), dim3(kBlock), 0, st, w.tay,
                     w.bes, F, two_pi * df * inv_c * rho, tc);
  auto kern = bi ? k_bwd_taylor<kTaylorK, kChebK, true> : k_bwd_taylor<kTaylorK, kChebK, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)rowblocks, (unsigned)ch), dim3(kBlock), 0, st, w.st,
                     w.pp, w.recd, w.recf, w.recg, w.groupc, sh.G, goff, (int)sh.ngroups, gpc, ant,
                     rx_off, w.mt, w.tay, w.bes, f0 * inv_c, df * inv_c, tc, rho, A, num_rx, F,
                     is_ti_radar);
  if (grad_meas)
    hipLaunchKernelGGL(k_bt_reduce_meas, dim3((A + kRedI - 1) / kRedI, (F + kRedJ - 1) / kRedJ),
                       dim3(kBlock), 0, st, (float2 *)grad_meas, w.st, (int)ch, A, F, scale);
  if (grad_ant)
    hipLaunchKernelGGL(k_bt_reduce_ant, dim3(mmn_blocks(3 * A, kBlock)), dim3(kBlock), 0, st,
                       grad_ant, w.pp, (int)ch, A, scale);
  return mmn_check_launch("mmn_bp_profile_backward");
}
