// Surface-normal vote kernels.
//
//   k_sn_mirror       the reference vote (surface_normal_est.cu:64-149) over a
//                     compacted voxel list: per row, the interpolated
//                     range-bin contribution c_i, weight Re(c_i conj(I))/|I|,
//                     and the weighted vector to the virtual antenna.  Same
//                     arithmetic and operation order as the reference.
//   k_sn_from_moment  completes the full-frequency vote from the moment the
//                     backprojection accumulates in the same pass:
//                     n = Re(conj(I) M) / |I|, M = sum_i vec_i c_i.
#include "mmn_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_sn_mirror(
    float *__restrict__ out, const int32_t *__restrict__ vidx, int64_t nvalid,
    const float2 *__restrict__ img, const float *__restrict__ xl,
    const float *__restrict__ yl, const float *__restrict__ zl,
    const float *__restrict__ ant, const float2 *__restrict__ meas,
    const float *__restrict__ rxo, float wavelength, float fft_spacing,
    int ny, int nz, int rows, int R, int S, int is_ti) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nvalid) return;
  const int64_t lin = vidx[t];
  float vx, vy, vz;
  voxel_coords(xl, yl, zl, ny, nz, lin, vx, vy, vz);
  const float2 I = img[lin];
  const float mag = sqrtf(I.x * I.x + I.y * I.y);

  float snx = 0.0f, sny = 0.0f, snz = 0.0f;
  for (int i = 0; i < rows; ++i) {
    const float px = ant[3 * i + 0], py = ant[3 * i + 1], pz = ant[3 * i + 2];
    const float ax = vx - px, ay = vy - py, az = vz - pz;
    const int rx = i % R;
    const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
    const float fwd = sqrtf(ax * ax + ay * ay + az * az);
    const float bx = ax - ox, by = ay - oy, bz = az - oz;
    const float back = sqrtf(bx * bx + by * by + bz * bz);
    float d = fwd + back;
    if (is_ti) d = (float)((double)d + 0.15);
    if (d < 0.0f || d > fft_spacing * (float)S) continue;
    const int bin = (int)floorf(d / fft_spacing / 2.0f);
    const float2 m = meas[(size_t)i * S + bin];
    const double th = -2.0 * kPi * (double)d / (double)wavelength;
    const float pr = (float)cos(th);
    const float pi = (float)sin(th);
    const float cr = m.x * pr - m.y * pi;
    const float ci = m.x * pi + m.y * pr;
    const float vecx = (px + ox / 2.0f) - vx;
    const float vecy = (py + oy / 2.0f) - vy;
    const float vecz = (pz + oz / 2.0f) - vz;
    float w = (cr * I.x) + (ci * I.y);
    w /= mag;
    snx += vecx * w;
    sny += vecy * w;
    snz += vecz * w;
  }
  out[3 * t + 0] = snx;
  out[3 * t + 1] = sny;
  out[3 * t + 2] = snz;
}

__global__ __launch_bounds__(kBlock) void k_sn_from_moment(
    float *__restrict__ out, const float2 *__restrict__ img,
    const float2 *__restrict__ mom, const int32_t *__restrict__ vidx,
    int64_t nvalid) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nvalid) return;
  const float2 I = img[vidx[t]];
  const float mag = sqrtf(I.x * I.x + I.y * I.y);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float2 m = mom[3 * t + k];
    out[3 * t + k] = fmaf(I.x, m.x, I.y * m.y) / mag;
  }
}

}  // namespace

extern "C" int mmn_cuda_sn_est(float *normals, const int32_t *vox_idx,
                               int64_t n_valid, const float *img,
                               const float *x_locs, const float *y_locs,
                               const float *z_locs, const float *ant,
                               const float *meas, const float *rx_off,
                               float slope, float wavelength,
                               float fft_spacing, int nx, int ny, int nz,
                               int num_rows, int num_rx, int samples,
                               int is_ti_radar, void *stream) {
  (void)slope;  // the interpolated vote uses no chirp slope
  MMN_REQUIRE(n_valid >= 0, MMN_ERR_ARG, "mmn_cuda_sn_est: n_valid < 0");
  if (n_valid == 0) return 0;
  MMN_REQUIRE(normals && vox_idx && img && x_locs && y_locs && z_locs && ant &&
                  meas && rx_off,
              MMN_ERR_NULL, "mmn_cuda_sn_est: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 &&
                  samples > 0,
              MMN_ERR_ARG, "mmn_cuda_sn_est: sizes must be positive");
  hipLaunchKernelGGL(k_sn_mirror, dim3(mmn_blocks(n_valid, kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, normals, vox_idx,
                     n_valid, (const float2 *)img, x_locs, y_locs, z_locs, ant,
                     (const float2 *)meas, rx_off, wavelength, fft_spacing, ny,
                     nz, num_rows, num_rx, samples, is_ti_radar);
  return mmn_check_launch("mmn_cuda_sn_est");
}

extern "C" int mmn_sn_from_moment(float *normals, const float *img,
                                  const float *moment, const int32_t *vox_idx,
                                  int64_t n_valid, void *stream) {
  MMN_REQUIRE(n_valid >= 0, MMN_ERR_ARG, "mmn_sn_from_moment: n_valid < 0");
  if (n_valid == 0) return 0;
  MMN_REQUIRE(normals && img && moment && vox_idx, MMN_ERR_NULL,
              "mmn_sn_from_moment: NULL pointer argument");
  hipLaunchKernelGGL(k_sn_from_moment, dim3(mmn_blocks(n_valid, kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, normals,
                     (const float2 *)img, (const float2 *)moment, vox_idx,
                     n_valid);
  return mmn_check_launch("mmn_sn_from_moment");
}
