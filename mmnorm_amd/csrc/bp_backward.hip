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
// Both reduce over voxels.  A workgroup of four waves owns one row, four
// blocks of K samples (one per wave) and one voxel chunk; the chunk is walked
// in steps of 256 voxels.  In each step every thread first forms one voxel's
// block-independent set-up -- the FP64 round-trip distance, the position
// gradient of the distance, the step phasor -- into LDS (double-buffered, one
// barrier per step), so that the four waves share it; then each wave runs its
// block over the step's voxels, two per lane at a time, with only its block's
// FP64-reduced anchor computed per wave.  Lanes keep lane-local sums, which
// one butterfly across the wave reduces at the end.  The block's K
// position-term coefficients are wave-uniform and come through scalar loads
// as SGPR operands.  Per (lane, voxel) the K conjugate phasors come from the
// anchor and a two-term FP32 recurrence; per sample the lane does the
// recurrence (2 FMA), one accumulate (2 add) and the position term (2 FMA).
// Each wave writes its partial sums to the workspace; a second kernel
// reduces the partials in a fixed order (FP64), so results are bitwise
// reproducible.
#include <algorithm>

#include "mmn_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kVV = 2;                 // voxels per lane at a time
constexpr int kStep = kBlock;          // voxels per workgroup step (one per thread)
constexpr int kPass = kWave * kVV;     // voxels of one wave pass over a step
constexpr int kPad = kStep;            // voxel records are padded to this
constexpr int kMinK = 16;              // smallest sample block (workspace sizing)
constexpr int kMaxK = 64;              // the position gradient alone (no accumulators)

// Per-voxel record: position and Re(g) in one 16-byte array element (rec4),
// Im(g) in a second array (recg), 20 bytes per voxel in all.  Records past
// the voxel count (up to a multiple of kPad) repeat the last voxel's
// position with g = 0.

int64_t pad_vox(int64_t nvox) { return (nvox + kPad - 1) / kPad * kPad; }

// Most voxel chunks a launch may use: at least ~64 voxels per chunk, at
// most 1024 chunks and ~1 GiB of partial sums.
int chunk_max(int A, int F, int64_t nvox) {
  int64_t m = (nvox + 63) / 64;
  const int64_t by_ws = ((int64_t)1 << 30) / ((int64_t)A * F * 8);
  if (m > by_ws) m = by_ws;
  if (m > 1024) m = 1024;
  return (int)(m < 1 ? 1 : m);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct BwdWs {
  float4 *rec4;   // [pad_vox]
  float *recg;    // [pad_vox]
  float2 *gs;     // [chunks][A][F]           adjoint partials
  float *gp;      // [chunks][nfbmax][A][3]   position partials
  float2 *coef;   // [A][nfb*K]               position-term coefficients
};

// Workspace of a launch with `sub` voxel chunks.
size_t bwd_layout(int A, int F, int64_t nvox, int sub, BwdWs *w, char *base) {
  const int64_t nv = pad_vox(nvox);
  const int nfbmax = (F + kMinK - 1) / kMinK;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  char *r4 = take(sizeof(float4) * (size_t)nv);
  char *rg = take(sizeof(float) * (size_t)nv);
  char *s = take(sizeof(float2) * (size_t)sub * A * F);
  char *g = take(sizeof(float) * (size_t)sub * nfbmax * A * 3);
  char *cf = take(sizeof(float2) * (size_t)A * (F + kMaxK));
  if (w) {
    w->rec4 = (float4 *)r4;
    w->recg = (float *)rg;
    w->gs = (float2 *)s;
    w->gp = (float *)g;
    w->coef = (float2 *)cf;
  }
  return off;
}

// Voxel chunks of a launch.  A workgroup walks every voxel of its chunk, and
// the workgroups of one chunk (all rows, all sample-block groups) run side
// by side, so a chunk whose records (20 bytes a voxel) fit well inside one
// XCD's 4 MB L2 is read from memory about once per XCD: chunks of at most
// kChunkRecBytes of records.  Measured at config 2 (1.57 M voxels, 3600
// rows, 256 samples), both gradients: 2 chunks (31 MB of records between
// them) fetched 27-30 GB per launch, 8 chunks 2.2 GB, 16 chunks 0.53 GB, 32
// chunks 0.69 GB, while the time went from 182.7 ms to 179.5-180.0 ms; the
// partial sums grow by A * F * 8 bytes a chunk (7.4 MB there).  Also at
// least 8 rounds of resident workgroups (occupancy x CUs), at least 256
// voxels a chunk and at most chunk_max.  The chunk length is a multiple of
// kPad; the count it implies never exceeds the planned one.
constexpr int64_t kChunkRecBytes = (int64_t)2 << 20;
constexpr int kRecBytes = 20;

struct BwdPlan {
  int ch = 1;         // chunks
  int64_t chunk = 0;  // voxels per chunk (multiple of kPad)
};

BwdPlan plan_bwd(int A, int F, int64_t nvox, int K, int64_t slots, int cap) {
  const int nfb = (F + K - 1) / K;
  const int nfbg = (nfb + kWavesPerBlock - 1) / kWavesPerBlock;
  const int64_t pairs = (int64_t)A * nfbg;
  const int64_t npad = pad_vox(nvox);
  int64_t ch = (8 * slots + pairs - 1) / pairs;
  ch = std::max(ch, (npad * kRecBytes + kChunkRecBytes - 1) / kChunkRecBytes);
  ch = std::min(ch, (nvox + 255) / 256);
  ch = std::min(ch, (int64_t)std::min(cap, chunk_max(A, F, nvox)));
  ch = std::max(ch, (int64_t)1);
  BwdPlan p;
  p.chunk = (npad + ch - 1) / ch;
  p.chunk = (p.chunk + kPad - 1) / kPad * kPad;
  p.ch = (int)((npad + p.chunk - 1) / p.chunk);
  return p;
}

__global__ __launch_bounds__(kBlock) void k_pack(
    float4 *__restrict__ rec4, float *__restrict__ recg, const float2 *__restrict__ g,
    const float *__restrict__ xl, const float *__restrict__ yl,
    const float *__restrict__ zl, const int32_t *__restrict__ vidx,
    int64_t nvox, int64_t npad, int ny, int nz) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= npad) return;
  const bool real = t < nvox;
  const int64_t src = real ? t : nvox - 1;
  const int64_t lin = vidx ? (int64_t)vidx[src] : src;
  float vx, vy, vz;
  voxel_coords(xl, yl, zl, ny, nz, lin, vx, vy, vz);
  const float2 gv = real ? g[t] : make_float2(0.0f, 0.0f);
  rec4[t] = make_float4(vx, vy, vz, gv.x);
  recg[t] = gv.y;
}

// Position-term coefficients (k_j Im s_ij, -k_j Re s_ij) of every (row,
// sample); rows hold nfb * K samples, zero past F.
__global__ __launch_bounds__(kBlock) void k_coef(float2 *__restrict__ coef,
                                                 const float2 *__restrict__ meas,
                                                 float k0, float dk, int F, int Fs,
                                                 int64_t AFs) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= AFs) return;
  const int a = (int)(t / Fs), j = (int)(t % Fs);
  float2 c = make_float2(0.0f, 0.0f);
  if (j < F) {
    const float2 sv = meas[(size_t)a * F + j];
    const float kj = fmaf((float)j, dk, k0);
    c = make_float2(kj * sv.y, -(kj * sv.x));
  }
  coef[t] = c;
}

// Per (row a, voxel v) and sample block the conjugate phasors
// z_t = g_v exp(+j 2 pi f_t d_av / c) for t = 0..K-1 satisfy, because the
// step exp(+j w) has modulus one, the real-coefficient recurrence
//     z_{t+1} = 2 cos(w) z_t - z_{t-1}
// (one FMA per component and step).  z_0 comes from the FP64-reduced anchor
// and z_1 = z_0 exp(+j w).  Rounding in this recurrence accumulates at most
// linearly, amplified by 1/sin(w); for the bands and ranges here
// (w = 2 pi df d / c between 0.05 and 1 rad) that stays below 1e-5 over a
// 32-sample block.
// Shared row-level set-up of one voxel: FP64 round-trip distance, the
// position gradient of the distance, the step phasor exp(+j w) and g_v.
struct VoxShared {
  double d;
  float gx, gy, gz, rr, ri, gr, gi, pad;
};

// One workgroup per (row a, group fbg of 4 sample blocks, voxel chunk c);
// wave w takes sample block 4 fbg + w (a wave past the row's last block runs
// along for the barriers and writes nothing).  MEAS: the adjoint (dL/dmeas)
// is accumulated; ANT: the position term.  With the adjoint the K = 32
// kernel fits in 128 VGPRs (64 of them the accumulators) at 4 waves per SIMD,
// the cap costing a few spills in the epilogue only; without it (position
// gradient only, as when the measurements take no gradient) there are no
// accumulators and it runs at 7 waves per SIMD.  Measured at config 2 (3600
// rows, 256 samples, 1.57 M voxels) on one box: 195 ms with both gradients,
// 143 ms for the position gradient alone, against 220 / 175 ms when every
// wave formed its own voxels' set-up.
template <int K, bool MEAS, bool ANT>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(MEAS ? 4 : 7))) void k_bwd(
    float2 *__restrict__ gs, float *__restrict__ gp,
    const float4 *__restrict__ rec4, const float *__restrict__ recg, int nvp,
    int chunk, const float *__restrict__ ant, const float2 *__restrict__ coef,
    const float *__restrict__ rxo, double f0c, float dfc32, double dfc, int A, int R,
    int F, int is_ti, int nfbmax, int nfb, int nfbg) {
  static_assert(K % 8 == 0, "coefficients are read 8 at a time");
  static_assert(MEAS || ANT, "nothing to compute");
  __shared__ VoxShared sh[2][kStep];
  const int lane = threadIdx.x % kWave;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t wg = blockIdx.x;
  const int a = (int)(wg % A);
  const int64_t rest = wg / A;
  const int fbg = (int)(rest % nfbg);
  const int c = (int)(rest / nfbg);
  const int fb = fbg * kWavesPerBlock + wv;
  const bool live = fb < nfb;
  const int fbl = live ? fb : nfb - 1;
  const int f_lo = fbl * K;
  const int nf = min(K, F - f_lo);
  const float2 *__restrict__ cf = coef + ((size_t)a * nfb + fbl) * K;

  const float px = ant[3 * a + 0], py = ant[3 * a + 1], pz = ant[3 * a + 2];
  const int rx = a % R;
  const float ox = rxo[3 * rx + 0], oy = rxo[3 * rx + 1], oz = rxo[3 * rx + 2];
  const bool mono = (ox == 0.0f) & (oy == 0.0f) & (oz == 0.0f);
  const double fc = fma((double)f_lo, dfc, f0c);
  const double d_off = is_ti ? 0.15 : 0.0;

  float accr[MEAS ? K : 1], acci[MEAS ? K : 1];
#pragma unroll
  for (int t = 0; t < (MEAS ? K : 1); ++t) accr[t] = acci[t] = 0.0f;
  float gpx = 0.0f, gpy = 0.0f, gpz = 0.0f;

  const int v0 = c * chunk;
  const int v1 = v0 + min(chunk, nvp - v0);
  int buf = 0;
  for (int vb0 = v0; vb0 < v1; vb0 += kStep, buf ^= 1) {
    {  // phase A: this thread's voxel
      const int v = vb0 + (int)threadIdx.x;
      const float4 p4 = rec4[v];
      const float gim = recg[v];
      const double dx = (double)p4.x - (double)px, dy = (double)p4.y - (double)py,
                   dz = (double)p4.z - (double)pz;
      const double s1 = fma(dz, dz, fma(dy, dy, dx * dx));
      const float y1 = __builtin_amdgcn_rsqf((float)s1);
      const double y1d = (double)y1;
      const double e0 = s1 * y1d;
      const double d1 = fma(0.5 * y1d, fma(-e0, e0, s1), e0);
      const float fx = p4.x - px, fy = p4.y - py, fz = p4.z - pz;
      double d;
      VoxShared o;
      if (mono) {
        d = d1 + d1;
        const float y2 = y1 + y1;
        o.gx = fx * y2;
        o.gy = fy * y2;
        o.gz = fz * y2;
      } else {
        const double ex = dx - (double)ox, ey = dy - (double)oy, ez = dz - (double)oz;
        d = d1 + sqrt_nr64(fma(ez, ez, fma(ey, ey, ex * ex)));
        const float bx = (float)ex, by = (float)ey, bz = (float)ez;
        const float n2 = rsqrtf(fmaf(bx, bx, fmaf(by, by, bz * bz)));
        o.gx = fmaf(bx, n2, fx * y1);
        o.gy = fmaf(by, n2, fy * y1);
        o.gz = fmaf(bz, n2, fz * y1);
      }
      d += d_off;
      const float ts = (float)d * dfc32;
      const float fr = ts - rintf(ts);
      o.rr = __builtin_amdgcn_cosf(fr);
      o.ri = __builtin_amdgcn_sinf(fr);
      o.gr = p4.w;
      o.gi = gim;
      o.d = d;
      o.pad = 0.0f;
      sh[buf][threadIdx.x] = o;
    }
    __syncthreads();
    // phase B: this wave's block over the step's voxels, two per lane at a time
#pragma unroll 1
    for (int j = 0; j < kStep / kPass; ++j) {
      float r0[kVV], i0[kVV], r1[kVV], i1[kVV], c2[kVV], tt[kVV], gx[kVV], gy[kVV], gz[kVV];
#pragma unroll
      for (int u = 0; u < kVV; ++u) {
        const VoxShared q = sh[buf][j * kPass + u * kWave + lane];
        float ar, ai;
        phasor_hw(q.d * fc, ar, ai);
        ai = -ai;
        r0[u] = q.gr * ar - q.gi * ai;
        i0[u] = q.gr * ai + q.gi * ar;
        r1[u] = r0[u] * q.rr - i0[u] * q.ri;
        i1[u] = r0[u] * q.ri + i0[u] * q.rr;
        c2[u] = q.rr + q.rr;
        gx[u] = q.gx;
        gy[u] = q.gy;
        gz[u] = q.gz;
        tt[u] = 0.0f;
      }
      float2 sp[8];
#pragma unroll
      for (int t = 0; t < K; ++t) {
        if (ANT && t % 8 == 0) {
          asm volatile("" ::: "memory");
#pragma unroll
          for (int qq = 0; qq < 8; ++qq) sp[qq] = cf[t + qq];
        }
        if constexpr (MEAS) {
          float sr = r0[0], si = i0[0];
#pragma unroll
          for (int u = 1; u < kVV; ++u) {
            sr += r0[u];
            si += i0[u];
          }
          accr[t] += sr;
          acci[t] += si;
        }
#pragma unroll
        for (int u = 0; u < kVV; ++u) {
          if constexpr (ANT) {
            tt[u] = fmaf(sp[t % 8].x, r0[u], tt[u]);
            tt[u] = fmaf(sp[t % 8].y, i0[u], tt[u]);
          }
          const float r2 = fmaf(c2[u], r1[u], -r0[u]);
          const float i2 = fmaf(c2[u], i1[u], -i0[u]);
          r0[u] = r1[u]; i0[u] = i1[u]; r1[u] = r2; i1[u] = i2;
        }
      }
      if constexpr (ANT) {
#pragma unroll
        for (int u = 0; u < kVV; ++u) {
          gpx = fmaf(-tt[u], gx[u], gpx);
          gpy = fmaf(-tt[u], gy[u], gpy);
          gpz = fmaf(-tt[u], gz[u], gpz);
        }
      }
    }
  }
  if constexpr (MEAS) {
    // Wave sums of the 2K adjoint values v = (Re acc_0..K-1, Im acc_0..K-1),
    // scattered: lane l ends with the total of v[l] (l < 2K).  Rounds over
    // partner lanes l ^ 32, l ^ 16, ..., l ^ 1, each adding own + partner as
    // a full butterfly does (so the totals are the butterfly's, bit for bit);
    // from o = K down, a lane keeps only the half of its values whose index
    // has bit o equal to its lane's and sends the other half, so the values a
    // lane holds halve each round (2K adds in all, not 12K, and no spills).
    constexpr int NV = 2 * K;
    float v[NV];
#pragma unroll
    for (int t = 0; t < K; ++t) {
      v[t] = accr[t];
      v[K + t] = acci[t];
    }
#pragma unroll
    for (int o = kWave / 2; o >= NV; o >>= 1) {
#pragma unroll
      for (int j = 0; j < NV; ++j) v[j] += __shfl_xor(v[j], o, kWave);
    }
#pragma unroll
    for (int o = NV / 2; o > 0; o >>= 1) {
      const bool up = (lane & o) != 0;
#pragma unroll
      for (int j = 0; j < o; ++j) {
        const float send = up ? v[j] : v[j + o];
        const float keep = up ? v[j + o] : v[j];
        v[j] = keep + __shfl_xor(send, o, kWave);
      }
    }
    if (live && lane < NV && lane % K < nf)
      reinterpret_cast<float *>(gs + ((size_t)c * A + a) * F + f_lo)[2 * (lane % K) + lane / K] =
          v[0];
  }
  if constexpr (ANT) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      gpx += __shfl_xor(gpx, o, kWave);
      gpy += __shfl_xor(gpy, o, kWave);
      gpz += __shfl_xor(gpz, o, kWave);
    }
  }
  if (live) {
    if (ANT && lane == 0) {
      float *__restrict__ o = gp + (((size_t)c * nfbmax + fb) * A + a) * 3;
      o[0] = gpx;
      o[1] = gpy;
      o[2] = gpz;
    }
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

// Resident workgroups of a backward kernel on the current device
// (occupancy x CUs); without a device (the workspace query may run on a host
// with no GPU), MI355X's 256 CUs and the kernel's register cap (4 workgroups
// per CU with the adjoint accumulators, 7 without).
template <int K, bool MEAS, bool ANT>
int64_t bwd_slots() {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bwd<K, MEAS, ANT>, kBlock, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = MEAS ? 4 : 7;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  (void)hipGetLastError();
  return (int64_t)per_cu * cus;
}

template <int K>
int bwd_chunks_any(int A, int F, int64_t nvox) {
  const int cap = 1 << 30;
  return std::max({plan_bwd(A, F, nvox, K, bwd_slots<K, true, true>(), cap).ch,
                   plan_bwd(A, F, nvox, K, bwd_slots<K, true, false>(), cap).ch,
                   plan_bwd(A, F, nvox, K, bwd_slots<K, false, true>(), cap).ch});
}

int bwd_chunks_ant64(int A, int F, int64_t nvox) {
  return plan_bwd(A, F, nvox, 64, bwd_slots<64, false, true>(), 1 << 30).ch;
}

}  // namespace

// Bytes of the largest workspace a launch with these sizes plans (over K = 16
// and 32 and the three output selections, and K = 64 for the position
// gradient alone).
extern "C" size_t mmn_bp_backward_workspace_size(int num_rows, int samples,
                                                 int64_t n_vox) {
  if (num_rows <= 0 || samples <= 0 || n_vox <= 0) return 0;
  const int sub = std::max({bwd_chunks_any<16>(num_rows, samples, n_vox),
                            bwd_chunks_any<32>(num_rows, samples, n_vox),
                            bwd_chunks_ant64(num_rows, samples, n_vox)});
  return bwd_layout(num_rows, samples, n_vox, sub, nullptr, nullptr);
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
  MMN_REQUIRE(K == 16 || K == 32 || (K == 64 && !grad_meas), MMN_ERR_ARG,
              "mmn_bp_backward: K must be 16 or 32, or 64 for the position gradient alone "
              "(got %d)", K);
  const int64_t nvox = vox_idx ? n_vox : (int64_t)nx * ny * nz;
  MMN_REQUIRE(nvox >= 0, MMN_ERR_ARG, "mmn_bp_backward: n_vox < 0");
  MMN_REQUIRE(pad_vox(nvox) <= INT32_MAX, MMN_ERR_ARG,
              "mmn_bp_backward: more than 2^31 - 129 voxels");
  hipStream_t st = (hipStream_t)stream;
  const int A = num_rows, F = samples;
  if (nvox == 0 || (!grad_meas && !grad_ant)) {
    int rc = 0;
    if (grad_meas)
      rc = mmn_zero_async(grad_meas, sizeof(float2) * (size_t)A * F, st, "mmn_bp_backward");
    if (rc == 0 && grad_ant)
      rc = mmn_zero_async(grad_ant, sizeof(float) * 3 * (size_t)A, st, "mmn_bp_backward");
    return rc;
  }
  MMN_REQUIRE(workspace, MMN_ERR_WS, "mmn_bp_backward: NULL workspace");
  const int64_t npad = pad_vox(nvox);
  const int nfb = (F + K - 1) / K;
  const int nfbmax = (F + kMinK - 1) / kMinK;
  // the gradients asked for select the kernel: both, the adjoint alone, or
  // the position gradient alone (no adjoint accumulators, twice the waves)
  using KernT = void (*)(float2 *, float *, const float4 *, const float *, int, int,
                         const float *, const float2 *, const float *, double, float, double,
                         int, int, int, int, int, int, int);
  const bool gm_on = grad_meas != nullptr, ga_on = grad_ant != nullptr;
  KernT kern;
  if (K == 64)
    kern = k_bwd<64, false, true>;
  else if (K == 16)
    kern = gm_on && ga_on ? k_bwd<16, true, true> : gm_on ? k_bwd<16, true, false>
                                                            : k_bwd<16, false, true>;
  else
    kern = gm_on && ga_on ? k_bwd<32, true, true> : gm_on ? k_bwd<32, true, false>
                                                            : k_bwd<32, false, true>;
  int per_cu = 0, dev = 0, cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipGetLastError();
  const int nfbg = (nfb + kWavesPerBlock - 1) / kWavesPerBlock;
  // the planned chunks, fewer if the workspace given holds fewer partials
  const int64_t slots = (int64_t)per_cu * cus;
  BwdPlan plan = plan_bwd(A, F, nvox, K, slots, 1 << 30);
  while (plan.ch > 1 && bwd_layout(A, F, nvox, plan.ch, nullptr, nullptr) > workspace_bytes)
    plan = plan_bwd(A, F, nvox, K, slots, plan.ch - 1);
  const size_t need = bwd_layout(A, F, nvox, plan.ch, nullptr, nullptr);
  MMN_REQUIRE(workspace_bytes >= need, MMN_ERR_WS,
              "mmn_bp_backward: workspace of %zu bytes needed", need);
  BwdWs w;
  bwd_layout(A, F, nvox, plan.ch, &w, (char *)workspace);
  const int ch = plan.ch;
  const int64_t chunk = plan.chunk;
  const int64_t nwgs = (int64_t)A * nfbg * ch;
  MMN_REQUIRE(nwgs * kBlock <= (int64_t)UINT32_MAX, MMN_ERR_ARG,
              "mmn_bp_backward: %lld workgroups of %d lanes exceed one launch (2^32 lanes); "
              "split the voxels or rows over calls", (long long)nwgs, kBlock);
  hipLaunchKernelGGL(k_pack, dim3(mmn_blocks(npad, kBlock)), dim3(kBlock), 0, st, w.rec4,
                     w.recg, (const float2 *)g, x_locs, y_locs, z_locs, vox_idx, nvox, npad,
                     ny, nz);
  const double inv_c = 1.0 / MMN_C0;
  const double two_pi = 6.283185307179586476925286766559;
  const float k0 = (float)(two_pi * f0 * inv_c), dk = (float)(two_pi * df * inv_c);
  const int Fs = nfb * K;
  hipLaunchKernelGGL(k_coef, dim3(mmn_blocks((int64_t)A * Fs, kBlock)), dim3(kBlock), 0, st,
                     w.coef, (const float2 *)meas, k0, dk, F, Fs, (int64_t)A * Fs);
  hipLaunchKernelGGL(kern, dim3((unsigned)nwgs), dim3(kBlock), 0, st, w.gs, w.gp, w.rec4,
                     w.recg, (int)npad, (int)chunk, ant, (const float2 *)w.coef, rx_off,
                     f0 * inv_c, (float)(df * inv_c), df * inv_c, A, num_rx, F, is_ti_radar,
                     nfbmax, nfb, nfbg);
  if (grad_meas)
    hipLaunchKernelGGL(k_reduce_meas, dim3(mmn_blocks((int64_t)A * F, kBlock)),
                       dim3(kBlock), 0, st, (float2 *)grad_meas, w.gs, (int)ch,
                       (int64_t)A * F, scale);
  if (grad_ant)
    hipLaunchKernelGGL(k_reduce_ant, dim3(mmn_blocks((int64_t)A * 3, kBlock)),
                       dim3(kBlock), 0, st, grad_ant, w.gp, (int)ch, nfb, nfbmax, A,
                       scale);
  return mmn_check_launch("mmn_bp_backward");
}
