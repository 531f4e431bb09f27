// Backward of the full-frequency backprojection (adjoint synthesis and the
// antenna-position gradient), recomputing phases instead of storing the
// rows x samples x voxels phasor tensor.
//
// Forward (per voxel v):  I_v = scale * sum_{i,j} s_ij phi_ijv,
//                         phi_ijv = exp(-j 2 pi f_j d_iv / c).
// With g_v = dL/dI_v (PyTorch convention) the backward is
//   dL/ds_ij = scale * sum_v g_v conj(phi_ijv)                    (adjoint)
//   dL/dp_i  = scale * sum_v Im( sum_j k_j s_ij conj(z_ijv) ) grad_p d_iv,
//              z_ijv = g_v conj(phi_ijv),  k_j = 2 pi f_j / c,
//              grad_p d = -(v-p)/|v-p| - (v-p-o)/|v-p-o|.
// Both reduce over voxels, so the mapping is transposed with respect to the
// forward: one lane per (row, block of K samples), all lanes of a wave walk
// the same voxel (wave-uniform voxel records through scalar loads).  Per
// (lane, voxel) the K conjugate phasors come from one FP64-reduced anchor and
// K-1 FP32 rotations; per sample the lane does one rotate (4 FMA), one
// accumulate (2 add) and the position term (2 FMA).  Each wave owns a voxel
// sub-range and writes its partial sums to the workspace; a second kernel
// reduces the partials in a fixed order (FP64), so results are bitwise
// reproducible.
#include <stdlib.h>

#include "mmn_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kTargetWaves = 8192;

struct VoxRec {  // 32 B, one scalar load per voxel
  float vx, vy, vz, pad0;
  float gr, gi, pad1, pad2;
};

struct BwdPlan {
  int sub;     // voxel sub-ranges (one per wave)
  int nfb;     // sample blocks of this launch
  int nfbmax;  // sample blocks the workspace is sized for (K = 8)
  int64_t chunk;
};

BwdPlan make_plan(int A, int F, int64_t nvox, int K) {
  BwdPlan p;
  const int agroups = (A + kWave - 1) / kWave;
  const int per_sub = agroups * ((F + 31) / 32);
  int64_t sub = (kTargetWaves + per_sub - 1) / per_sub;
  const int64_t by_vox = (nvox + 63) / 64;  // at least ~64 voxels per wave
  if (sub > by_vox) sub = by_vox;
  if (sub < 1) sub = 1;
  if (sub > 1024) sub = 1024;
  sub = (sub + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock;
  p.sub = (int)sub;
  p.nfb = K > 0 ? (F + K - 1) / K : 0;
  p.nfbmax = (F + 7) / 8;
  p.chunk = (nvox + p.sub - 1) / p.sub;
  return p;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return (v && *v) ? atoi(v) : dflt;
}

struct BwdWs {
  VoxRec *rec;    // [nvox]
  float2 *gs;     // [sub][A][F]
  float *gp;      // [sub][nfbmax][A][3]
};

size_t bwd_layout(int A, int F, int64_t nvox, BwdWs *w, char *base) {
  const BwdPlan p = make_plan(A, F, nvox, 8);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  char *r = take(sizeof(VoxRec) * (size_t)nvox);
  char *s = take(sizeof(float2) * (size_t)p.sub * A * F);
  char *g = take(sizeof(float) * (size_t)p.sub * p.nfbmax * A * 3);
  if (w) {
    w->rec = (VoxRec *)r;
    w->gs = (float2 *)s;
    w->gp = (float *)g;
  }
  return off;
}

__global__ __launch_bounds__(kBlock) void k_pack(
    VoxRec *__restrict__ rec, const float2 *__restrict__ g,
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    int64_t nvox, int ny, int nz) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nvox) return;
  const int64_t lin = vidx ? (int64_t)vidx[t] : t;
  VoxRec r;
  voxel_coords(xl, yl, zl, ny, nz, lin, r.vx, r.vy, r.vz);
  const float2 gv = g[t];
  r.gr = gv.x;
  r.gi = gv.y;
  r.pad0 = r.pad1 = r.pad2 = 0.0f;
  rec[t] = r;
}

// Per (lane row a, voxel v) state of one sample block: the conjugate phasors
// z_t = g_v exp(+j 2 pi f_t d_av / c) for t = 0..K-1 satisfy, because the
// step exp(+j w) has modulus one, the real-coefficient recurrence
//     z_{t+1} = 2 cos(w) z_t - z_{t-1}
// (one FMA per component and step).  z_0 comes from the FP64-reduced anchor
// and z_1 = z_0 exp(+j w).  Rounding in this recurrence accumulates at most
// linearly, amplified by 1/sin(w); for the bands and ranges here
// (w = 2 pi df d / c between 0.05 and 1 rad) that stays below 1e-5 over a
// 32-sample block.
struct VoxStart {
  float z0r, z0i, z1r, z1i, c2;  // c2 = 2 cos(w)
  float gx, gy, gz;              // grad_p d
};

__device__ __forceinline__ VoxStart vox_start(const VoxRec &r, float px, float py,
                                              float pz, float ox, float oy, float oz,
                                              bool mono, int is_ti, double fc,
                                              double dfc, float weight) {
  VoxStart v;
  double d = roundtrip_dist64(r.vx, r.vy, r.vz, px, py, pz, ox, oy, oz, mono);
  if (is_ti) d += 0.15;
  float ar, ai, rr, ri;
  phasor_hw(d * fc, ar, ai);  // exp(-j ...); conjugated below
  phasor_hw(d * dfc, rr, ri);
  ai = -ai;
  ri = -ri;
  const float gr = r.gr * weight, gi = r.gi * weight;
  v.z0r = gr * ar - gi * ai;
  v.z0i = gr * ai + gi * ar;
  v.z1r = v.z0r * rr - v.z0i * ri;
  v.z1i = v.z0r * ri + v.z0i * rr;
  v.c2 = rr + rr;
  const float ax = r.vx - px, ay = r.vy - py, az = r.vz - pz;
  const float n1 = rsqrtf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
  float gx = ax * n1, gy = ay * n1, gz = az * n1;
  if (mono) {
    gx += gx;
    gy += gy;
    gz += gz;
  } else {
    const float bx = ax - ox, by = ay - oy, bz = az - oz;
    const float n2 = rsqrtf(fmaf(bx, bx, fmaf(by, by, bz * bz)));
    gx = fmaf(bx, n2, gx);
    gy = fmaf(by, n2, gy);
    gz = fmaf(bz, n2, gz);
  }
  // d(d)/dp = -(unit vectors): the sign is applied with the position term.
  v.gx = gx;
  v.gy = gy;
  v.gz = gz;
  return v;
}

template <int K, int VV>
__global__ __launch_bounds__(kBlock) void k_bwd(
    float2 *__restrict__ gs, float *__restrict__ gp,
    const VoxRec *__restrict__ rec, int64_t nvox, int64_t chunk,
    const float *__restrict__ ant, const float2 *__restrict__ meas,
    const float *__restrict__ rxo, double f0c, double dfc, float k0,
    float dk, int A, int R, int F, int is_ti, int nfbmax) {
  const int lane = threadIdx.x % kWave;
  const int sub = blockIdx.z * kWavesPerBlock + threadIdx.x / kWave;
  const int fb = blockIdx.y;
  const int a = min(blockIdx.x * kWave + lane, A - 1);
  const bool live_row = (blockIdx.x * kWave + lane) < A;
  const int f_lo = fb * K;
  const int nf = min(K, F - f_lo);

  const float px = ant[3 * a + 0], py = ant[3 * a + 1], pz = ant[3 * a + 2];
  const int rx = a % R;
  const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
  const bool mono = (ox == 0.0f) & (oy == 0.0f) & (oz == 0.0f);

  // k_j * s_ij for this lane's samples (zero past the end of the row).
  float spr[K], spi[K], accr[K], acci[K];
  const float2 *__restrict__ row = meas + (size_t)a * F + f_lo;
#pragma unroll
  for (int t = 0; t < K; ++t) {
    float2 s = make_float2(0.0f, 0.0f);
    if (t < nf) s = row[t];
    const float kj = fmaf((float)(f_lo + t), dk, k0);
    spr[t] = kj * s.x;
    spi[t] = kj * s.y;
    accr[t] = acci[t] = 0.0f;
  }
  float gpx = 0.0f, gpy = 0.0f, gpz = 0.0f;
  const double fc = fma((double)f_lo, dfc, f0c);

  const int64_t v0 = (int64_t)sub * chunk;
  const int64_t v1 = min(nvox, v0 + chunk);
  // VV voxels per iteration: VV independent recurrences per lane.  A short
  // last group repeats its final voxel with zero weight.
  // Voxel records are read with vector loads (every lane the same address),
  // whose in-order completion lets the next group's records be requested one
  // iteration ahead: the wait for them then falls after a full group of work.
  // `zero` is a lane value the compiler cannot see through, which keeps these
  // uniform loads on the vector path instead of scalar loads (whose waits
  // cover every outstanding scalar load and so cannot be overlapped).
  int zero;
  asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
  const VoxRec *__restrict__ rv = rec + zero;
  VoxRec nxt[VV];
  if (v0 < v1) {
#pragma unroll
    for (int u = 0; u < VV; ++u) nxt[u] = rv[min(v0 + u, v1 - 1)];
  }
  for (int64_t v = v0; v < v1; v += VV) {
    VoxRec cur[VV];
#pragma unroll
    for (int u = 0; u < VV; ++u) cur[u] = nxt[u];
    if (v + VV < v1) {
#pragma unroll
      for (int u = 0; u < VV; ++u) nxt[u] = rv[min(v + VV + u, v1 - 1)];
    }
    VoxStart p[VV];
    float w[VV];
#pragma unroll
    for (int u = 0; u < VV; ++u) {
      w[u] = (v + u < v1) ? 1.0f : 0.0f;
      p[u] = vox_start(cur[u], px, py, pz, ox, oy, oz, mono, is_ti, fc, dfc, w[u]);
    }
    float r0[VV], i0[VV], r1[VV], i1[VV], t1[VV], t2[VV];
#pragma unroll
    for (int u = 0; u < VV; ++u) {
      r0[u] = p[u].z0r; i0[u] = p[u].z0i; r1[u] = p[u].z1r; i1[u] = p[u].z1i;
      t1[u] = t2[u] = 0.0f;
    }
#pragma unroll
    for (int t = 0; t < K; ++t) {
      float sr = r0[0], si = i0[0];
#pragma unroll
      for (int u = 1; u < VV; ++u) {
        sr += r0[u];
        si += i0[u];
      }
      accr[t] += sr;
      acci[t] += si;
#pragma unroll
      for (int u = 0; u < VV; ++u) {
        t1[u] = fmaf(spi[t], r0[u], t1[u]);
        t2[u] = fmaf(spr[t], i0[u], t2[u]);
        const float r2 = fmaf(p[u].c2, r1[u], -r0[u]);
        const float i2 = fmaf(p[u].c2, i1[u], -i0[u]);
        r0[u] = r1[u]; i0[u] = i1[u]; r1[u] = r2; i1[u] = i2;
      }
    }
#pragma unroll
    for (int u = 0; u < VV; ++u) {
      const float tt = t1[u] - t2[u];
      gpx = fmaf(-tt, p[u].gx, gpx);
      gpy = fmaf(-tt, p[u].gy, gpy);
      gpz = fmaf(-tt, p[u].gz, gpz);
    }
  }
  if (live_row) {
    float2 *__restrict__ out = gs + ((size_t)sub * A + a) * F + f_lo;
#pragma unroll
    for (int t = 0; t < K; ++t)
      if (t < nf) out[t] = make_float2(accr[t], acci[t]);
    float *__restrict__ o = gp + (((size_t)sub * nfbmax + fb) * A + a) * 3;
    o[0] = gpx;
    o[1] = gpy;
    o[2] = gpz;
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce_meas(
    float2 *__restrict__ grad, const float2 *__restrict__ gs, int sub,
    int64_t AF, double scale) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= AF) return;
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < sub; ++s) {
    const float2 v = gs[(size_t)s * AF + t];
    sr += v.x;
    si += v.y;
  }
  grad[t] = make_float2((float)(scale * sr), (float)(scale * si));
}

__global__ __launch_bounds__(kBlock) void k_reduce_ant(
    float *__restrict__ grad, const float *__restrict__ gp, int sub, int nfb,
    int nfbmax, int A, double scale) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= A * 3) return;
  double s = 0.0;
  for (int q = 0; q < sub; ++q)
    for (int b = 0; b < nfb; ++b) s += gp[((size_t)q * nfbmax + b) * A * 3 + t];
  grad[t] = (float)(scale * s);
}

}  // namespace

extern "C" size_t mmn_bp_backward_workspace_size(int num_rows, int samples,
                                                 int64_t n_vox) {
  if (num_rows <= 0 || samples <= 0 || n_vox <= 0) return 0;
  return bwd_layout(num_rows, samples, n_vox, nullptr, nullptr);
}

extern "C" int mmn_bp_backward(float *grad_meas, float *grad_ant,
                               const float *g, double scale,
                               const float *x_locs, const float *y_locs,
                               const float *z_locs, const int32_t *vox_idx,
                               int64_t n_vox, const float *ant,
                               const float *meas, const float *rx_off,
                               double f0, double df, int nx, int ny, int nz,
                               int num_rows, int num_rx, int samples,
                               int is_ti_radar, int K, void *workspace,
                               size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(g && x_locs && y_locs && z_locs && ant && meas && rx_off,
              MMN_ERR_NULL, "mmn_bp_backward: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && num_rows > 0 && num_rx > 0 &&
                  samples > 0,
              MMN_ERR_ARG, "mmn_bp_backward: sizes must be positive");
  if (K == 0) K = 32;
  MMN_REQUIRE(K == 8 || K == 16 || K == 32 || K == 48, MMN_ERR_ARG,
              "mmn_bp_backward: K must be 8, 16, 32 or 48 (got %d)", K);
  const int64_t nvox = vox_idx ? n_vox : (int64_t)nx * ny * nz;
  MMN_REQUIRE(nvox >= 0, MMN_ERR_ARG, "mmn_bp_backward: n_vox < 0");
  hipStream_t st = (hipStream_t)stream;
  const int A = num_rows, F = samples;
  if (nvox == 0 || (!grad_meas && !grad_ant)) {
    hipError_t e = hipSuccess;
    if (grad_meas) e = hipMemsetAsync(grad_meas, 0, sizeof(float2) * (size_t)A * F, st);
    if (e == hipSuccess && grad_ant) e = hipMemsetAsync(grad_ant, 0, sizeof(float) * 3 * (size_t)A, st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_bp_backward: %s", hipGetErrorString(e));
      return (int)e;
    }
    return 0;
  }
  const size_t need = bwd_layout(A, F, nvox, nullptr, nullptr);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_bp_backward: workspace of %zu bytes needed", need);
  BwdWs w;
  bwd_layout(A, F, nvox, &w, (char *)workspace);
  const BwdPlan p = make_plan(A, F, nvox, K);

  hipLaunchKernelGGL(k_pack, dim3(mmn_blocks(nvox, kBlock)), dim3(kBlock), 0,
                     st, w.rec, (const float2 *)g, x_locs, y_locs, z_locs,
                     vox_idx, nvox, ny, nz);
  const double inv_c = 1.0 / MMN_C0;
  const double two_pi = 6.283185307179586476925286766559;
  const float k0 = (float)(two_pi * f0 * inv_c);
  const float dk = (float)(two_pi * df * inv_c);
  const dim3 grid((A + kWave - 1) / kWave, p.nfb, p.sub / kWavesPerBlock);
#define MMN_BWD_LAUNCH(KK, VV)                                                \
  hipLaunchKernelGGL((k_bwd<KK, VV>), grid, dim3(kBlock), 0, st, w.gs, w.gp, w.rec, \
                     nvox, p.chunk, ant, (const float2 *)meas, rx_off,        \
                     f0 * inv_c, df * inv_c, k0, dk, A, num_rx,               \
                     F, is_ti_radar, p.nfbmax)
  const int vv = env_int("MMN_BWD_VV", 2);
  if (K == 8) MMN_BWD_LAUNCH(8, 2);
  else if (K == 16) { if (vv == 1) MMN_BWD_LAUNCH(16, 1); else MMN_BWD_LAUNCH(16, 2); }
  else if (K == 48) { if (vv == 2) MMN_BWD_LAUNCH(48, 2); else MMN_BWD_LAUNCH(48, 1); }
  else { if (vv == 1) MMN_BWD_LAUNCH(32, 1); else MMN_BWD_LAUNCH(32, 2); }
#undef MMN_BWD_LAUNCH
  if (grad_meas)
    hipLaunchKernelGGL(k_reduce_meas, dim3(mmn_blocks((int64_t)A * F, kBlock)),
                       dim3(kBlock), 0, st, (float2 *)grad_meas, w.gs, p.sub,
                       (int64_t)A * F, scale);
  if (grad_ant)
    hipLaunchKernelGGL(k_reduce_ant, dim3(mmn_blocks((int64_t)A * 3, kBlock)),
                       dim3(kBlock), 0, st, grad_ant, w.gp, p.sub, p.nfb,
                       p.nfbmax, A, scale);
  return mmn_check_launch("mmn_bp_backward");
}
