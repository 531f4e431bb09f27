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
constexpr int kRowC = 12;

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
  c[10] = 0.0;
  c[11] = 0.0;
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
    const float *__restrict__ ant, const float *__restrict__ rxo,
    const double *__restrict__ rowc, const float4 *__restrict__ coef, float thf, int M,
    int rows, int R, int rps) {
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
  auto fetch = [&](int i, float (&f)[VPT], float4 (&a)[VPT], float4 (&b)[VPT]) {
    const int rx = i % R;
    const bool mono = (rxo[3 * rx + 0] == 0.0f) & (rxo[3 * rx + 1] == 0.0f) &
                      (rxo[3 * rx + 2] == 0.0f);
    const double *__restrict__ c = rowc + (size_t)kRowC * i;
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
  auto finish = [&](int i, const float (&f)[VPT], const float4 (&a)[VPT],
                    const float4 (&b)[VPT], float (&Sr)[VPT], float (&Si)[VPT]) {
    const int rx = i % R;
    const float vax = ant[3 * i + 0] + rxo[3 * rx + 0] / 2.0f;
    const float vay = ant[3 * i + 1] + rxo[3 * rx + 1] / 2.0f;
    const float vaz = ant[3 * i + 2] + rxo[3 * rx + 2] / 2.0f;
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
    float f[VPT];
    float4 a[VPT], b[VPT];
    fetch(i, f, a, b);
    finish(i, f, a, b, Sr, Si);
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
// Backward of the range-profile forward, for a real loss with g = dL/dI.
//   dL/dp_i = scale sum_v Im(conj(g_v) H_i(d_iv)) grad_p d_iv,
//   H_i(d)  = sum_j k_j s_ij exp(-j 2 pi f_j d / c)   (k_j = 2 pi f_j / c),
// with H_i interpolated from its own profile table exactly as the forward
// interpolates P_i; and the adjoint of the forward's interpolation (the
// exact gradient of the profile forward): each voxel spreads
//   w_k(f) exp(+j 2 pi thf (f + 1 - k)) g_v,   k = 0..3,
// onto table points m - 1 + k of its row (w_k the cubic's Lagrange weights
// at nodes -1, 0, 1, 2, thf = f_c h / c the carrier's turns per interval),
// and then
//   dL/ds_ij = scale sum_m G_i[m] exp(+j 2 pi f_j (dlo_i + m h) / c).
// k_bwd_prof gives each wave one row and a chunk of voxels: the row's table
// points accumulate in the wave's own LDS grid (ds_add, no other wave
// touches it) and the position term in lanes, so nothing crosses waves until
// the per-(chunk, row) partials are reduced in a fixed order.
// --------------------------------------------------------------------------
constexpr int kBwdMaxM = 2048;  // LDS grid per wave: M float2 (4 waves: 64 KiB)
constexpr int kBwdPad = 128;    // voxel records padded to 2 per lane
constexpr int kDftK = 32;       // grid points per FP64 re-anchor in k_bwd_prof_dft

// One leg as leg64, also returning the FP32 inverse length.
__device__ __forceinline__ double leg64y(double vx, double vy, double vz, double vv,
                                         const double *__restrict__ c, float &inv) {
  const double s = fma(vz, c[2], fma(vy, c[1], fma(vx, c[0], vv + c[-1])));
  const float y = __builtin_amdgcn_rsqf((float)s);
  inv = y;
  const double yd = (double)y;
  const double d0 = s * yd;
  return fma(0.5 * yd, fma(-d0, d0, s), d0);
}

__global__ __launch_bounds__(kBlock) void k_bwdp_pack(
    float4 *__restrict__ rec4, float *__restrict__ recg, const float2 *__restrict__ g,
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx, int64_t n,
    int64_t npad, int ny, int nz) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= npad) return;
  const bool real = t < n;
  const int64_t src = real ? t : n - 1;
  const int64_t lin = vidx ? (int64_t)vidx[src] : src;
  float vx, vy, vz;
  voxel_coords(xl, yl, zl, ny, nz, lin, vx, vy, vz);
  const float2 gv = real ? g[t] : make_float2(0.0f, 0.0f);
  rec4[t] = make_float4(vx, vy, vz, gv.x);
  recg[t] = gv.y;
}

struct TapPhases {
  float r[4], i[4];  // exp(+j 2 pi thf (1 - k)), k = 0..3
};

__global__ __launch_bounds__(kBlock) void k_bwd_prof(
    float2 *__restrict__ gpart, float *__restrict__ ppart, const float4 *__restrict__ rec4,
    const float *__restrict__ recg, int64_t npad, int64_t chunk,
    const float *__restrict__ ant, const float *__restrict__ rxo,
    const double *__restrict__ rowc, const float4 *__restrict__ coefh, float thf,
    TapPhases tp, int M, int A, int R) {
  extern __shared__ float2 lds_grid[];  // [4][M]
  const int lane = threadIdx.x % 64;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
  const int i_raw = blockIdx.x * 4 + wv;
  const bool live = i_raw < A;
  const int i = live ? i_raw : A - 1;
  const int c = blockIdx.y;
  float2 *__restrict__ grid = lds_grid + (size_t)wv * M;
  for (int m = lane; m < M; m += 64) grid[m] = make_float2(0.0f, 0.0f);
  __syncthreads();

  const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
  const int rx = i % R;
  const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
  const bool mono = (ox == 0.0f) & (oy == 0.0f) & (oz == 0.0f);
  const float qx = px + ox, qy = py + oy, qz = pz + oz;
  const double *__restrict__ rc = rowc + (size_t)kRowC * i;
  const float4 *__restrict__ crow = coefh + (size_t)i * M * 2;
  const float umax = (float)(M - 3);
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;

  // chunk and npad are multiples of kBwdPad, so vb + 64 is a record too
  const int64_t v0 = (int64_t)c * chunk;
  const int64_t v1 = min(npad, v0 + chunk);
  for (int64_t vb = v0 + lane; vb < v1; vb += kBwdPad) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const float4 r = rec4[vb + 64 * k];
      const float gr = r.w, gi = recg[vb + 64 * k];
      const double dx = r.x, dy = r.y, dz = r.z;
      const double vv = fma(dz, dz, fma(dy, dy, dx * dx));
      float y1, y2 = 0.0f;
      double ds = leg64y(dx, dy, dz, vv, rc + 2, y1);
      if (!mono) ds += leg64y(dx, dy, dz, vv, rc + 6, y2);
      const double u = fma(ds, rc[9], rc[0]);
      const int m = (int)fminf(fmaxf((float)u, 1.0f), umax);
      const float f = __builtin_amdgcn_fmed3f((float)(u - (double)m), -1.0f, 2.0f);
      // H at the voxel: the cubic times exp(-j 2 pi thf f)
      const float4 a = crow[2 * m], b = crow[2 * m + 1];
      const float qr = fmaf(fmaf(fmaf(b.z, f, b.x), f, a.z), f, a.x);
      const float qi = fmaf(fmaf(fmaf(b.w, f, b.y), f, a.w), f, a.y);
      const float ph = thf * f;
      const float cs = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
      const float hr = fmaf(qr, cs, qi * sn), hi = fmaf(qi, cs, -(qr * sn));
      // position term: Im(conj(g) H) grad_p d
      const float tt = fmaf(gr, hi, -(gi * hr));
      const float ax = px - r.x, ay = py - r.y, az = pz - r.z;
      if (mono) {
        const float w1 = 2.0f * tt * y1;
        gx = fmaf(w1, ax, gx);
        gy = fmaf(w1, ay, gy);
        gz = fmaf(w1, az, gz);
      } else {
        const float w1 = tt * y1, w2 = tt * y2;
        gx = fmaf(w1, ax, fmaf(w2, qx - r.x, gx));
        gy = fmaf(w1, ay, fmaf(w2, qy - r.y, gy));
        gz = fmaf(w1, az, fmaf(w2, qz - r.z, gz));
      }
      // adjoint: g exp(+j 2 pi thf f), spread with the Lagrange weights
      const float er = fmaf(gr, cs, -(gi * sn)), ei = fmaf(gr, sn, gi * cs);
      const float fp1 = f + 1.0f, fm1 = f - 1.0f, fm2 = f - 2.0f;
      const float ff1 = f * fm1;
      float w[4];
      w[0] = -(ff1 * fm2) * (1.0f / 6.0f);
      w[1] = (fp1 * fm1 * fm2) * 0.5f;
      w[2] = -(fp1 * f * fm2) * 0.5f;
      w[3] = (fp1 * ff1) * (1.0f / 6.0f);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float zr = w[q] * tp.r[q], zi = w[q] * tp.i[q];
        atomicAdd(&grid[m - 1 + q].x, fmaf(zr, er, -(zi * ei)));
        atomicAdd(&grid[m - 1 + q].y, fmaf(zr, ei, zi * er));
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    gx += __shfl_xor(gx, o, 64);
    gy += __shfl_xor(gy, o, 64);
    gz += __shfl_xor(gz, o, 64);
  }
  __syncthreads();
  if (live) {
    float2 *__restrict__ out = gpart + ((size_t)c * A + i) * M;
    for (int m = lane; m < M; m += 64) out[m] = grid[m];
    if (lane == 0) {
      float *__restrict__ o = ppart + ((size_t)c * A + i) * 3;
      o[0] = gx;
      o[1] = gy;
      o[2] = gz;
    }
  }
}

// Fixed-order FP64 sum of the chunks' grids.
__global__ __launch_bounds__(kBlock) void k_bwd_prof_merge(float2 *__restrict__ graw,
                                                           const float2 *__restrict__ gpart,
                                                           int chunks, int64_t AM) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= AM) return;
  double a = 0.0, b = 0.0;
  for (int q = 0; q < chunks; ++q) {
    const float2 v = gpart[(size_t)q * AM + t];
    a += v.x;
    b += v.y;
  }
  graw[t] = make_float2((float)a, (float)b);
}

// dL/ds_ij = scale sum_m G_i[m] exp(+j 2 pi f_j (dlo_i + m h) / c): one lane
// per sample j of row blockIdx.y; the row's grid values are wave-uniform
// (scalar loads).  Horner in exp(+j 2 pi f_j h / c) from the top, re-anchored
// every kDftK points by an FP64-reduced phasor.
__global__ __launch_bounds__(kBlock) void k_bwd_prof_dft(
    float2 *__restrict__ grad, const float2 *__restrict__ graw,
    const double *__restrict__ dlo, double h, double f0c, double dfc, int M, int F,
    double scale) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * kBlock + threadIdx.x;
  const double fj = fma((double)min(j, F - 1), dfc, f0c);  // revolutions per metre
  float rr, ri;
  phasor_quarter(-4.0 * fj * h, rr, ri);  // exp(+j 2 pi f_j h / c)
  const float2 *__restrict__ gi = graw + (size_t)i * M;
  const double d0 = dlo[i];
  float cr = 0.0f, ci = 0.0f;
  for (int b0 = 0; b0 < M; b0 += kDftK) {
    const int len = min(kDftK, M - b0);
    const float2 top = gi[b0 + len - 1];
    float hr = top.x, hi = top.y;
    for (int t = len - 2; t >= 0; --t) {
      const float2 s = gi[b0 + t];
      const float tr = fmaf(-hi, ri, fmaf(hr, rr, s.x));
      const float ti = fmaf(hi, rr, fmaf(hr, ri, s.y));
      hr = tr;
      hi = ti;
    }
    float ar, ai;
    phasor_hw(-fj * fma((double)b0, h, d0), ar, ai);  // exp(+j 2 pi f_j d_b0 / c)
    cr = fmaf(ar, hr, fmaf(-ai, hi, cr));
    ci = fmaf(ar, hi, fmaf(ai, hr, ci));
  }
  if (j < F)
    grad[(size_t)i * F + j] = make_float2((float)(scale * cr), (float)(scale * ci));
}

__global__ __launch_bounds__(kBlock) void k_bwd_prof_ant(float *__restrict__ grad,
                                                         const float *__restrict__ ppart,
                                                         int chunks, int A, double scale) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= 3 * A) return;
  double s = 0.0;
  for (int q = 0; q < chunks; ++q) s += ppart[(size_t)q * A * 3 + t];
  grad[t] = (float)(scale * s);
}

struct BwdPWs {
  double *dlo, *rowc;
  float2 *tab;
  float4 *coef;
  float4 *rec4;
  float *recg;
  float2 *gpart, *graw;
  float *ppart;
};

int64_t bwdp_pad(int64_t n) { return (n + kBwdPad - 1) / kBwdPad * kBwdPad; }

// Most voxel chunks a launch may use: ~1 GiB of partial grids, at most 64.
int bwdp_chunks_max(int A, int M) {
  int64_t c = ((int64_t)1 << 30) / ((int64_t)A * M * 8);
  if (c > 64) c = 64;
  return c < 1 ? 1 : (int)c;
}

size_t bwdp_layout(int A, int M, int64_t n, BwdPWs *w, char *base) {
  const int64_t np = bwdp_pad(n);
  const int cm = bwdp_chunks_max(A, M);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  char *d = take(sizeof(double) * (size_t)A);
  char *rc = take(sizeof(double) * kRowC * (size_t)A);
  char *t = take(sizeof(float2) * (size_t)A * M);
  char *cf = take(sizeof(float4) * 2 * (size_t)A * M);
  char *r4 = take(sizeof(float4) * (size_t)np);
  char *rg = take(sizeof(float) * (size_t)np);
  char *gp = take(sizeof(float2) * (size_t)cm * A * M);
  char *gr = take(sizeof(float2) * (size_t)A * M);
  char *pp = take(sizeof(float) * 3 * (size_t)cm * A);
  if (w) {
    w->dlo = (double *)d;
    w->rowc = (double *)rc;
    w->tab = (float2 *)t;
    w->coef = (float4 *)cf;
    w->rec4 = (float4 *)r4;
    w->recg = (float *)rg;
    w->gpart = (float2 *)gp;
    w->graw = (float2 *)gr;
    w->ppart = (float *)pp;
  }
  return off;
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
                     y_locs, z_locs, vox_idx, n, nx, ny, nz, ant, rx_off, w.rowc,
                     (const float4 *)w.coef, (float)(fcc * g.h), M, A, num_rx, rps);
  if (nseg > 1) mmn_rs_reduce(st, (float2 *)img, (float2 *)moment, im, mo, n, nseg);
  return mmn_check_launch("mmn_bp_profile");
}

extern "C" size_t mmn_bp_profile_backward_workspace_size(int64_t n_vox, int num_rows,
                                                         int samples, double f0, double df,
                                                         const double *bbox) {
  ProfGeom g;
  if (n_vox <= 0 || num_rows <= 0 || samples <= 0 || !bbox ||
      !prof_geom(samples, f0, df, bbox, &g) || g.M > kBwdMaxM)
    return 0;
  return bwdp_layout(num_rows, g.M, n_vox, nullptr, nullptr);
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
  ProfGeom geo;
  MMN_REQUIRE(prof_geom(samples, f0, df, bbox, &geo), MMN_ERR_ARG,
              "mmn_bp_profile_backward: f0, df and the bounding box must be finite (box min "
              "<= max), and the box's round-trip span at most %d table points", kMaxM);
  MMN_REQUIRE(geo.M <= kBwdMaxM, MMN_ERR_ARG,
              "mmn_bp_profile_backward: %d table points per row exceed the %d an LDS grid "
              "holds (a smaller box or mmn_bp_backward)", geo.M, kBwdMaxM);
  int64_t n = (int64_t)nx * ny * nz;
  if (vox_idx) {
    MMN_REQUIRE(n_vox >= 0, MMN_ERR_ARG, "mmn_bp_profile_backward: n_vox < 0");
    n = n_vox;
  }
  const int A = num_rows, F = samples, M = geo.M;
  hipStream_t st = (hipStream_t)stream;
  if (n == 0 || (!grad_meas && !grad_ant)) {
    hipError_t e = hipSuccess;
    if (grad_meas) e = hipMemsetAsync(grad_meas, 0, sizeof(float2) * (size_t)A * F, st);
    if (e == hipSuccess && grad_ant)
      e = hipMemsetAsync(grad_ant, 0, sizeof(float) * 3 * (size_t)A, st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_bp_profile_backward: %s", hipGetErrorString(e));
      return (int)e;
    }
    return 0;
  }
  const size_t need = bwdp_layout(A, M, n, nullptr, nullptr);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_bp_profile_backward: workspace of %zu bytes needed", need);
  BwdPWs w;
  bwdp_layout(A, M, n, &w, (char *)workspace);
  Box box;
  for (int k = 0; k < 6; ++k) box.b[k] = bbox[k];
  const double inv_c = 1.0 / MMN_C0;
  const double two_pi = 6.283185307179586476925286766559;
  const double tc = 0.5 * (double)(F - 1);
  const double fcc = (f0 + tc * df) * inv_c;
  hipLaunchKernelGGL(k_prof_rows, dim3(mmn_blocks(A, kBlock)), dim3(kBlock), 0, st, w.dlo,
                     w.rowc, ant, rx_off, box, A, num_rx, is_ti_radar, geo.h);
  const int64_t np = bwdp_pad(n);
  if (grad_ant) {
    hipLaunchKernelGGL(k_prof_table, dim3(mmn_blocks(M, kBlock), A), dim3(kBlock), 0, st,
                       w.tab, (const float2 *)meas, w.dlo, geo.h, df * inv_c, tc, M, F, 1,
                       (float)(two_pi * f0 * inv_c), (float)(two_pi * df * inv_c));
    hipLaunchKernelGGL(k_prof_coef, dim3(mmn_blocks((int64_t)A * M, kBlock)), dim3(kBlock), 0,
                       st, w.coef, w.tab, w.dlo, geo.h, fcc, M, (int64_t)A * M);
  } else {
    // the position term is not wanted: a zero table makes it vanish
    hipError_t e = hipMemsetAsync(w.coef, 0, sizeof(float4) * 2 * (size_t)A * M, st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_bp_profile_backward: %s", hipGetErrorString(e));
      return (int)e;
    }
  }
  hipLaunchKernelGGL(k_bwdp_pack, dim3(mmn_blocks(np, kBlock)), dim3(kBlock), 0, st, w.rec4,
                     w.recg, (const float2 *)g, x_locs, y_locs, z_locs, vox_idx, n, np, ny,
                     nz);
  // voxel chunks: enough workgroups for eight rounds of resident slots, at
  // least 1024 voxels per chunk, at most the workspace's count
  const size_t lds = sizeof(float2) * 4 * (size_t)M;
  int per_cu = 0, dev = 0, cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bwd_prof, kBlock, lds) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipGetLastError();
  const int64_t groups = (A + 3) / 4;
  int64_t ch = (8 * (int64_t)per_cu * cus + groups - 1) / groups;
  const int forced = env_int("MMN_PROFB_CHUNKS", 0);
  if (forced > 0) ch = forced;
  if (ch > (np + 1023) / 1024) ch = (np + 1023) / 1024;
  if (ch > bwdp_chunks_max(A, M)) ch = bwdp_chunks_max(A, M);
  if (ch < 1) ch = 1;
  int64_t chunk = (np + ch - 1) / ch;
  chunk = (chunk + kBwdPad - 1) / kBwdPad * kBwdPad;
  ch = (np + chunk - 1) / chunk;
  TapPhases tp;
  const double thf = fcc * geo.h;
  for (int k = 0; k < 4; ++k) {
    double ph = thf * (double)(1 - k);
    ph -= rint(ph);
    tp.r[k] = (float)cos(two_pi * ph);
    tp.i[k] = (float)sin(two_pi * ph);
  }
  hipLaunchKernelGGL(k_bwd_prof, dim3((unsigned)groups, (unsigned)ch), dim3(kBlock), lds, st,
                     w.gpart, w.ppart, w.rec4, w.recg, np, chunk, ant, rx_off, w.rowc,
                     (const float4 *)w.coef, (float)thf, tp, M, A, num_rx);
  if (grad_meas) {
    hipLaunchKernelGGL(k_bwd_prof_merge, dim3(mmn_blocks((int64_t)A * M, kBlock)),
                       dim3(kBlock), 0, st, w.graw, w.gpart, (int)ch, (int64_t)A * M);
    hipLaunchKernelGGL(k_bwd_prof_dft, dim3(mmn_blocks(F, kBlock), A), dim3(kBlock), 0, st,
                       (float2 *)grad_meas, w.graw, w.dlo, geo.h, f0 * inv_c, df * inv_c, M, F,
                       scale);
  }
  if (grad_ant)
    hipLaunchKernelGGL(k_bwd_prof_ant, dim3(mmn_blocks(3 * A, kBlock)), dim3(kBlock), 0, st,
                       grad_ant, w.ppart, (int)ch, A, scale);
  return mmn_check_launch("mmn_bp_profile_backward");
}
