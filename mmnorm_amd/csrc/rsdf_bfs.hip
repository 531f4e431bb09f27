// RSDF breadth-first search, FIFO-exact, level-synchronous on the GPU.
//
// The reference (sdf.py:180-222) runs one FIFO BFS per component.  FIFO order
// is level order, and inside level L+1 it is the order in which level-L
// voxels discovered their children: a child belongs to the first frontier
// position p (then the first direction d) that reaches it.  So for every
// level:
//   claim  each frontier position p offers key = 6p + d to every enterable
//          neighbour; the neighbour keeps the minimum (atomicMin on unique
//          keys: the winner does not depend on scheduling),
//   count  each p finds which of its neighbours it won (bit mask) and each
//          block sums its winners,
//   scan   one block turns the block sums into block offsets and writes the
//          next frontier size,
//   emit   each p writes its winners, in direction order, at its offset: the
//          next frontier comes out sorted by key, i.e. in FIFO order, and
//          every child's rsdf uses its winning parent:
//              rsdf[c] = rsdf[p] + dot(dir * dx, normals[c])  (sdf.py:222).
// Components never share voxels (labels differ), so all components run in
// one frontier, seeded in the reference's component order; each keeps its
// own FIFO order because keys only compare voxels of the same component.
//
// Each block owns one contiguous chunk of the frontier, so block order is
// position order.  The frontier size lives in device memory; the host
// enqueues levels in batches and reads the size back once per batch.
#include "mmn_internal.h"

namespace {

constexpr int kGrid = 1024;
constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kLevelsPerBatch = 16;

__constant__ int kDX[6] = {-1, 0, 1, 0, 0, 0};
__constant__ int kDY[6] = {0, 1, 0, -1, 0, 0};
__constant__ int kDZ[6] = {0, 0, 0, 0, 1, -1};

struct Ws {
  int32_t *claim;   // [V]  minimum key offered this level
  int32_t *front0;  // [V]
  int32_t *front1;  // [V]
  uint8_t *pmask;   // [V]  winners of frontier position p (bit d)
  int32_t *bcnt;    // [kGrid]
  int32_t *boff;    // [kGrid]
  int32_t *nf;      // [2]  frontier sizes by level parity
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t ws_layout(int64_t V, Ws *w, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  char *c = take(sizeof(int32_t) * V);
  char *f0 = take(sizeof(int32_t) * V);
  char *f1 = take(sizeof(int32_t) * V);
  char *pm = take(V);
  char *bc = take(sizeof(int32_t) * kGrid);
  char *bo = take(sizeof(int32_t) * kGrid);
  char *nf = take(sizeof(int32_t) * 2);
  if (w) {
    w->claim = (int32_t *)c;
    w->front0 = (int32_t *)f0;
    w->front1 = (int32_t *)f1;
    w->pmask = (uint8_t *)pm;
    w->bcnt = (int32_t *)bc;
    w->boff = (int32_t *)bo;
    w->nf = (int32_t *)nf;
  }
  return off;
}

__device__ __forceinline__ void chunk_of(int n, int &start, int &end) {
  const int chunk = (n + kGrid - 1) / kGrid;
  start = min(n, (int)blockIdx.x * chunk);
  end = min(n, start + chunk);
}

// Linear index of neighbour d of (x, y, z), or -1 when outside the grid.
__device__ __forceinline__ int64_t neighbour(int x, int y, int z, int d,
                                             int nx, int ny, int nz) {
  const int ax = x + kDX[d], ay = y + kDY[d], az = z + kDZ[d];
  if (ax < 0 || ay < 0 || az < 0 || ax >= nx || ay >= ny || az >= nz) return -1;
  return ((int64_t)ax * ny + ay) * nz + az;
}

__device__ __forceinline__ void decode(int64_t u, int ny, int nz, int &x,
                                       int &y, int &z) {
  const int64_t yz = (int64_t)ny * nz;
  x = (int)(u / yz);
  const int64_t r = u - (int64_t)x * yz;
  y = (int)(r / nz);
  z = (int)(r - (int64_t)y * nz);
}

__global__ void k_init(double *__restrict__ rsdf, int32_t *__restrict__ hops,
                       int32_t *__restrict__ claim, int64_t V) {
  const double qnan = __longlong_as_double(0x7ff8000000000000LL);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
       i += (int64_t)gridDim.x * blockDim.x) {
    rsdf[i] = qnan;
    hops[i] = -1;
    claim[i] = INT32_MAX;
  }
}

__global__ void k_seed(double *__restrict__ rsdf, int32_t *__restrict__ hops,
                       int32_t *__restrict__ front, int32_t *__restrict__ nf,
                       const int32_t *__restrict__ seeds, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const int32_t s = seeds[t];
    front[t] = s;
    hops[s] = 0;
    rsdf[s] = 0.0;
  }
  if (t == 0) {
    nf[0] = n;
    nf[1] = 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_claim(
    const int32_t *__restrict__ front, const int32_t *__restrict__ nfp,
    int32_t *__restrict__ claim, const int32_t *__restrict__ hops,
    const double *__restrict__ nrm, const int32_t *__restrict__ lab, int nx,
    int ny, int nz) {
  int start, end;
  chunk_of(*nfp, start, end);
  for (int p = start + threadIdx.x; p < end; p += kBlock) {
    const int32_t u = front[p];
    int x, y, z;
    decode(u, ny, nz, x, y, z);
    const int32_t lu = lab[u];
    for (int d = 0; d < 6; ++d) {
      const int64_t c = neighbour(x, y, z, d, nx, ny, nz);
      if (c < 0 || hops[c] != -1) continue;
      if (isnan(nrm[3 * c]) || lab[c] != lu) continue;
      atomicMin(&claim[c], 6 * p + d);
    }
  }
}

__device__ __forceinline__ int block_sum(int v, int *sh) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) sh[w] = v;
  __syncthreads();
  int tot = 0;
  for (int i = 0; i < kBlock / kWave; ++i) tot += sh[i];
  __syncthreads();
  return tot;
}

__global__ __launch_bounds__(kBlock) void k_count(
    const int32_t *__restrict__ front, const int32_t *__restrict__ nfp,
    const int32_t *__restrict__ claim, const int32_t *__restrict__ hops,
    uint8_t *__restrict__ pmask, int32_t *__restrict__ bcnt, int nx, int ny,
    int nz) {
  __shared__ int sh[kBlock / kWave];
  int start, end;
  chunk_of(*nfp, start, end);
  int cnt = 0;
  for (int p = start + threadIdx.x; p < end; p += kBlock) {
    const int32_t u = front[p];
    int x, y, z;
    decode(u, ny, nz, x, y, z);
    unsigned m = 0;
    for (int d = 0; d < 6; ++d) {
      const int64_t c = neighbour(x, y, z, d, nx, ny, nz);
      if (c < 0 || hops[c] != -1) continue;
      if (claim[c] == 6 * p + d) m |= 1u << d;
    }
    pmask[p] = (uint8_t)m;
    cnt += __popc(m);
  }
  const int tot = block_sum(cnt, sh);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// One block of kGrid threads: exclusive scan of the per-block counts.
__global__ __launch_bounds__(kGrid) void k_scan(const int32_t *__restrict__ bcnt,
                                                int32_t *__restrict__ boff,
                                                int32_t *__restrict__ nf_next) {
  __shared__ int sh[kGrid / kWave];
  const int t = threadIdx.x;
  const int v = bcnt[t];
  int inc = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const int n = __shfl_up(inc, o, kWave);
    if ((t % kWave) >= o) inc += n;
  }
  const int w = t / kWave;
  if ((t % kWave) == kWave - 1) sh[w] = inc;
  __syncthreads();
  int before = 0;
  for (int i = 0; i < w; ++i) before += sh[i];
  boff[t] = before + inc - v;
  if (t == kGrid - 1) *nf_next = before + inc;
}

__global__ __launch_bounds__(kBlock) void k_emit(
    const int32_t *__restrict__ front, const int32_t *__restrict__ nfp,
    const uint8_t *__restrict__ pmask, const int32_t *__restrict__ boff,
    int32_t *__restrict__ next, int32_t *__restrict__ hops,
    double *__restrict__ rsdf, const double *__restrict__ nrm, double dx,
    int level, int nx, int ny, int nz) {
#pragma clang fp contract(off)
  __shared__ int sh[kBlock / kWave];
  int start, end;
  chunk_of(*nfp, start, end);
  int base = boff[blockIdx.x];
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  for (int r0 = start; r0 < end; r0 += kBlock) {
    const int p = r0 + threadIdx.x;
    const unsigned m = (p < end) ? pmask[p] : 0u;
    const int c = __popc(m);
    int inc = c;
    for (int o = 1; o < kWave; o <<= 1) {
      const int n = __shfl_up(inc, o, kWave);
      if (l >= o) inc += n;
    }
    if (l == kWave - 1) sh[w] = inc;
    __syncthreads();
    int before = 0, round_tot = 0;
    for (int i = 0; i < kBlock / kWave; ++i) {
      if (i < w) before += sh[i];
      round_tot += sh[i];
    }
    __syncthreads();
    if (m) {
      int out = base + before + inc - c;
      const int32_t u = front[p];
      int x, y, z;
      decode(u, ny, nz, x, y, z);
      const double prev = rsdf[u];
      for (int d = 0; d < 6; ++d) {
        if (!(m & (1u << d))) continue;
        const int64_t ch = neighbour(x, y, z, d, nx, ny, nz);
        next[out++] = (int32_t)ch;
        hops[ch] = level + 1;
        // dot(dir * dx, n) as the reference forms it: three products, summed.
        const double t0 = ((double)kDX[d] * dx) * nrm[3 * ch + 0];
        const double t1 = ((double)kDY[d] * dx) * nrm[3 * ch + 1];
        const double t2 = ((double)kDZ[d] * dx) * nrm[3 * ch + 2];
        rsdf[ch] = prev + ((t0 + t1) + t2);
      }
    }
    base += round_tot;
  }
}

}  // namespace

extern "C" size_t mmn_rsdf_bfs_workspace_size(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return ws_layout((int64_t)nx * ny * nz, nullptr, nullptr);
}

extern "C" int mmn_rsdf_bfs(double *rsdf, int32_t *hops, const double *normals,
                            const int32_t *labels, const int32_t *seeds,
                            int n_seeds, double dx, int nx, int ny, int nz,
                            void *workspace, size_t workspace_bytes,
                            void *stream) {
  MMN_REQUIRE(rsdf && hops && normals && labels, MMN_ERR_NULL,
              "mmn_rsdf_bfs: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0 && n_seeds >= 0, MMN_ERR_ARG,
              "mmn_rsdf_bfs: sizes must be positive");
  const int64_t V = (int64_t)nx * ny * nz;
  MMN_REQUIRE(V <= (int64_t)(INT32_MAX / 6), MMN_ERR_ARG,
              "mmn_rsdf_bfs: grid of %lld voxels exceeds the int32 key range",
              (long long)V);
  MMN_REQUIRE(n_seeds == 0 || seeds, MMN_ERR_NULL, "mmn_rsdf_bfs: seeds NULL");
  MMN_REQUIRE(n_seeds <= V, MMN_ERR_ARG, "mmn_rsdf_bfs: more seeds than voxels");
  const size_t need = ws_layout(V, nullptr, nullptr);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_rsdf_bfs: workspace of %zu bytes needed", need);
  Ws w;
  ws_layout(V, &w, (char *)workspace);
  hipStream_t st = (hipStream_t)stream;

  hipLaunchKernelGGL(k_init, dim3(kGrid), dim3(kBlock), 0, st, rsdf, hops,
                     w.claim, V);
  hipLaunchKernelGGL(k_seed, dim3(mmn_blocks(n_seeds > 0 ? n_seeds : 1, kBlock)),
                     dim3(kBlock), 0, st, rsdf, hops, w.front0, w.nf, seeds,
                     n_seeds);
  int rc = mmn_check_launch("mmn_rsdf_bfs(init)");
  if (rc) return rc;
  if (n_seeds == 0) return 0;

  int32_t *fr[2] = {w.front0, w.front1};
  for (int level = 0; level <= V;) {
    for (int b = 0; b < kLevelsPerBatch; ++b, ++level) {
      const int cur = level & 1, nxt = cur ^ 1;
      hipLaunchKernelGGL(k_claim, dim3(kGrid), dim3(kBlock), 0, st, fr[cur],
                         w.nf + cur, w.claim, hops, normals, labels, nx, ny, nz);
      hipLaunchKernelGGL(k_count, dim3(kGrid), dim3(kBlock), 0, st, fr[cur],
                         w.nf + cur, w.claim, hops, w.pmask, w.bcnt, nx, ny, nz);
      hipLaunchKernelGGL(k_scan, dim3(1), dim3(kGrid), 0, st, w.bcnt, w.boff,
                         w.nf + nxt);
      hipLaunchKernelGGL(k_emit, dim3(kGrid), dim3(kBlock), 0, st, fr[cur],
                         w.nf + cur, w.pmask, w.boff, fr[nxt], hops, rsdf,
                         normals, dx, level, nx, ny, nz);
    }
    rc = mmn_check_launch("mmn_rsdf_bfs(level)");
    if (rc) return rc;
    int32_t n_next = 0;
    hipError_t e = hipMemcpyAsync(&n_next, w.nf + (level & 1), sizeof(int32_t),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_rsdf_bfs: %s", hipGetErrorString(e));
      return (int)e;
    }
    if (n_next == 0) break;
  }
  return 0;
}
