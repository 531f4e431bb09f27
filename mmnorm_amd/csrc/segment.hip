// Binary-image segmentation before the RSDF search (sdf.py:33-120): the Li
// threshold of |image|^exp and 6-connected component labelling with the
// per-label sizes the reference's size filter uses.
//
// Li threshold (Li & Tam's iterative minimum cross-entropy, the iteration of
// skimage.filters.threshold_li that sdf.py:60-70 calls): with x = a - min(a),
//     t <- (m_b - m_f) / (ln m_b - ln m_f),
//     m_f = mean{x > t},  m_b = mean{x <= t},
// from t = mean(x), until t stops changing, or m_b = 0.  Here a = |z|^exp is
// formed in FP64 and every sum is an FP64 reduction in a fixed order, so the
// result is deterministic; the state lives on the device and the host reads
// it back once per batch of iterations.
//
// Labelling: lock-free union-find over the 6-neighbourhood.  Each foreground
// voxel unites with its foreground neighbours at x-1, y-1, z-1; a union hooks
// the larger root under the smaller with a compare-and-swap, so parents only
// ever decrease and the final root of a component is its smallest linear
// index.  The unions run in two launches: inside 8 x 8 x 16 tiles, in LDS
// (k_ccl_local, which also writes every parent), then across tile faces on
// the tiles' roots in global memory (k_ccl_border) -- so a large component
// merges per tile rather than per voxel at its one root.  Labels are then
// numbered 1..n by that index, i.e. in the order a C-order raster scan first
// meets each component -- scipy.ndimage.label's numbering.
#include "mmn_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kGrid = 1024;           // blocks of the fixed-order reductions
constexpr int kItersPerBatch = 16;
constexpr int kScanChunk = kBlock * 16;  // voxels per block of the label scan

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------ Li --
struct LiState {
  double lo, hi;         // min and max of a
  double t_curr, t_next; // thresholds on x = a - lo
  int done, iters, max_iter, pad;
};

struct LiWs {
  double *a;        // [n]  |z|^exp
  double *part;     // [kGrid][4]  fixed-order partial sums
  LiState *st;
};

size_t li_layout(int64_t n, LiWs *w, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  char *a = take(sizeof(double) * (size_t)n);
  char *p = take(sizeof(double) * kGrid * 4);
  char *s = take(sizeof(LiState));
  if (w) {
    w->a = (double *)a;
    w->part = (double *)p;
    w->st = (LiState *)s;
  }
  return off;
}

// Sum over the block with a shared array; thread 0 gets the total.
template <typename T, int N>
__device__ __forceinline__ void block_sum(T (&v)[N], T (*sh)[kBlock]) {
#pragma unroll
  for (int q = 0; q < N; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < N; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = sh[q][0];
}

// a = |z|^exp; per-block min, max and sum of a.
__global__ __launch_bounds__(kBlock) void k_li_prep(double *__restrict__ a,
                                                    double *__restrict__ part,
                                                    const float2 *__restrict__ z,
                                                    int64_t n, double ex) {
  __shared__ double sh[3][kBlock];
  double lo = INFINITY, hi = -INFINITY, sum = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)kGrid * kBlock) {
    const float2 v = z[i];
    const double m = hypot((double)v.x, (double)v.y);
    const double p = pow(m, ex);
    a[i] = p;
    lo = fmin(lo, p);
    hi = fmax(hi, p);
    sum += p;
  }
  sh[0][threadIdx.x] = lo;
  sh[1][threadIdx.x] = hi;
  sh[2][threadIdx.x] = sum;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
      sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x + 0] = sh[0][0];
    part[4 * blockIdx.x + 1] = sh[1][0];
    part[4 * blockIdx.x + 2] = sh[2][0];
  }
}

// One block: min, max, initial guess t = mean(a - lo) = mean(a) - lo.
__global__ __launch_bounds__(kBlock) void k_li_start(LiState *__restrict__ st,
                                                     const double *__restrict__ part,
                                                     int64_t n, int max_iter) {
  __shared__ double sh[3][kBlock];
  double lo = INFINITY, hi = -INFINITY, sum = 0.0;
  for (int b = threadIdx.x; b < kGrid; b += kBlock) {
    lo = fmin(lo, part[4 * b + 0]);
    hi = fmax(hi, part[4 * b + 1]);
  }
  // the sum in a fixed order: thread-strided, then the tree below
  for (int b = threadIdx.x; b < kGrid; b += kBlock) sum += part[4 * b + 2];
  sh[0][threadIdx.x] = lo;
  sh[1][threadIdx.x] = hi;
  sh[2][threadIdx.x] = sum;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
      sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    LiState s;
    s.lo = sh[0][0];
    s.hi = sh[1][0];
    s.t_next = sh[2][0] / (double)n - s.lo;
    s.t_curr = -1.0;  // differs from any t_next >= 0: the loop runs once at least
    s.iters = 0;
    s.max_iter = max_iter;
    s.pad = 0;
    // a constant image is its own threshold (no iteration)
    s.done = s.hi == s.lo;
    if (s.done) s.t_next = 0.0;
    *st = s;
  }
}

// Sums at t = t_next (the iteration's new t_curr): foreground x > t.
__global__ __launch_bounds__(kBlock) void k_li_sums(double *__restrict__ part,
                                                    const double *__restrict__ a,
                                                    const LiState *__restrict__ st,
                                                    int64_t n) {
  if (st->done) return;
  __shared__ double sh[4][kBlock];
  const double lo = st->lo, t = st->t_next;
  double v[4] = {0.0, 0.0, 0.0, 0.0};  // sum_f, cnt_f, sum_b, cnt_b
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)kGrid * kBlock) {
    const double x = a[i] - lo;
    if (x > t) {
      v[0] += x;
      v[1] += 1.0;
    } else {
      v[2] += x;
      v[3] += 1.0;
    }
  }
  block_sum<double, 4>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) part[4 * blockIdx.x + q] = v[q];
  }
}

// One block: the update t_curr <- t_next, t_next <- Li step, stop tests.
__global__ __launch_bounds__(kBlock) void k_li_update(LiState *__restrict__ st,
                                                      const double *__restrict__ part) {
  if (st->done) return;
  __shared__ double sh[4][kBlock];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < kGrid; b += kBlock) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] += part[4 * b + q];
  }
  block_sum<double, 4>(v, sh);
  if (threadIdx.x == 0) {
    LiState s = *st;
    s.t_curr = s.t_next;
    const double mf = v[0] / v[1];  // NaN for an empty foreground, as np.mean
    const double mb = v[2] / v[3];
    if (mb == 0.0) {
      s.done = 1;  // log(0): keep t_next
    } else {
      s.t_next = (mb - mf) / (log(mb) - log(mf));
      s.iters += 1;  // iterations = threshold updates
      // stop when t no longer moves (the partition is stable), on NaN, or
      // at the iteration cap
      s.done = !(fabs(s.t_next - s.t_curr) > 0.0) || s.iters >= s.max_iter;
    }
    *st = s;
  }
}

// below[i] = a[i] < threshold, with the same a the iteration used.
__global__ __launch_bounds__(kBlock) void k_li_mask(uint8_t *__restrict__ below,
                                                    const double *__restrict__ a,
                                                    int64_t n, double thr) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) below[i] = a[i] < thr ? 1 : 0;
}

// ----------------------------------------------------------- labelling --
struct CclWs {
  int32_t *par;    // [V]  union-find parents (foreground), -1 background
  int32_t *bsum;   // [nchunks]  roots per scan chunk, then chunk offsets
  int32_t *count;  // [1]
};

int64_t ccl_chunks(int64_t V) { return (V + kScanChunk - 1) / kScanChunk; }

size_t ccl_layout(int64_t V, CclWs *w, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  char *p = take(sizeof(int32_t) * (size_t)V);
  char *b = take(sizeof(int32_t) * (size_t)ccl_chunks(V));
  char *c = take(sizeof(int32_t));
  if (w) {
    w->par = (int32_t *)p;
    w->bsum = (int32_t *)b;
    w->count = (int32_t *)c;
  }
  return off;
}

__device__ __forceinline__ int32_t load_par(const int32_t *par, int32_t x) {
  // device-coherent read: parents are rewritten by other workgroups
  return __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving.  Every parent is an ancestor with a smaller
// index, and parents only decrease, so a stale read is still an ancestor and
// the walk ends at a root (or at a node that was a root when read: unite
// re-checks with its compare-and-swap).
__device__ int32_t find_root(int32_t *par, int32_t x) {
  while (true) {
    const int32_t p = load_par(par, x);
    if (p == x) return x;
    const int32_t gp = load_par(par, p);
    if (gp == p) return p;
    __hip_atomic_store(par + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gp;
  }
}

// Root of x without writes (used once every union is done: a halving store
// racing with another thread's final store could leave a non-root parent).
__device__ int32_t root_of(const int32_t *par, int32_t x) {
  int32_t p = load_par(par, x);
  while (p != x) {
    x = p;
    p = load_par(par, x);
  }
  return x;
}

__device__ void unite(int32_t *par, int32_t a, int32_t b) {
  int32_t ra = find_root(par, a), rb = find_root(par, b);
  while (ra != rb) {
    if (ra < rb) {
      const int32_t t = ra;
      ra = rb;
      rb = t;
    }
    // hook the larger root under the smaller one
    int32_t expect = ra;
    if (__hip_atomic_compare_exchange_strong(par + ra, &expect, rb, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
    // ra gained a parent meanwhile: climb on from it
    ra = find_root(par, expect);
    rb = find_root(par, rb);
  }
}

// Tiles of 8 x 8 x 16 voxels, one thread per voxel (1024); local index
// t = (lx * 8 + ly) * 16 + lz orders a tile's voxels as their linear indices.
// Larger tiles leave fewer face voxels to unite globally: on a 1.5M-voxel
// field with one large component the border unions took 179 / 118 / 93 us
// with 8 x 8 x 4 / 8 x 8 x 8 / 8 x 8 x 16 tiles, a union per face voxel
// (32 us at 8 x 8 x 16 with the repeated pairs skipped, k_ccl_border).
constexpr int kTX = 8, kTY = 8, kTZ = 16;
constexpr int kTile = kTX * kTY * kTZ;

struct TileVoxel {
  int lx, ly, lz, x, y, z;
  int64_t v;
  bool in;
};

__device__ __forceinline__ TileVoxel tile_voxel(int nx, int ny, int nz) {
  const int ntz = (nz + kTZ - 1) / kTZ, nty = (ny + kTY - 1) / kTY;
  const int64_t b = blockIdx.x;
  const int tz = (int)(b % ntz), ty = (int)((b / ntz) % nty), tx = (int)(b / ((int64_t)ntz * nty));
  TileVoxel r;
  const int t = threadIdx.x;
  r.lx = t / (kTY * kTZ);
  r.ly = (t / kTZ) % kTY;
  r.lz = t % kTZ;
  r.x = tx * kTX + r.lx;
  r.y = ty * kTY + r.ly;
  r.z = tz * kTZ + r.lz;
  r.in = r.x < nx && r.y < ny && r.z < nz;
  r.v = r.in ? ((int64_t)r.x * ny + r.y) * nz + r.z : 0;
  return r;
}

// The union-find of unite / find_root on a tile's LDS parents (workgroup
// scope): parents only decrease, roots are a component's smallest index.
__device__ int32_t lfind(int32_t *lp, int32_t x) {
  while (true) {
    const int32_t p = __hip_atomic_load(lp + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (p == x) return x;
    const int32_t gp = __hip_atomic_load(lp + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (gp == p) return p;
    __hip_atomic_store(lp + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = gp;
  }
}

__device__ void lunite(int32_t *lp, int32_t a, int32_t b) {
  int32_t ra = lfind(lp, a), rb = lfind(lp, b);
  while (ra != rb) {
    if (ra < rb) {
      const int32_t t = ra;
      ra = rb;
      rb = t;
    }
    int32_t expect = ra;
    if (__hip_atomic_compare_exchange_strong(lp + ra, &expect, rb, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
      return;
    ra = lfind(lp, expect);
    rb = lfind(lp, rb);
  }
}

// Components inside each tile, in LDS; every voxel's parent becomes its tile
// component's root (the smallest linear index of that component within the
// tile, a root: its own parent), background -1.  Writes every voxel of the
// grid.
__global__ __launch_bounds__(kTile) void k_ccl_local(int32_t *__restrict__ par,
                                                      const uint8_t *__restrict__ mask,
                                                      int nx, int ny, int nz) {
  __shared__ int32_t lp[kTile];
  __shared__ uint8_t lf[kTile];
  const TileVoxel tv = tile_voxel(nx, ny, nz);
  const int t = threadIdx.x;
  const bool fg = tv.in && mask[tv.v];
  lp[t] = fg ? t : -1;
  lf[t] = fg;
  __syncthreads();
  if (fg) {
    if (tv.lz > 0 && lf[t - 1]) lunite(lp, t, t - 1);
    if (tv.ly > 0 && lf[t - kTZ]) lunite(lp, t, t - kTZ);
    if (tv.lx > 0 && lf[t - kTY * kTZ]) lunite(lp, t, t - kTY * kTZ);
  }
  __syncthreads();
  if (!tv.in) return;
  int32_t g = -1;
  if (fg) {
    int32_t r = t;
    for (int32_t q = lp[r]; q != r; q = lp[r]) r = q;
    const int rx = tv.x - tv.lx + r / (kTY * kTZ), ry = tv.y - tv.ly + (r / kTZ) % kTY,
              rz = tv.z - tv.lz + r % kTZ;
    g = (int32_t)(((int64_t)rx * ny + ry) * nz + rz);
  }
  par[tv.v] = g;
}

// One union per active lane across a tile face, on the parents of its two
// voxels (ancestors in the same sets).  The face voxels of a tile mostly
// join the same two tile components, so a lane whose pair of parents equals
// the previous active lane's in its wave skips its union: that lane joins
// the same two sets.  On a field with one large component this removes most
// of the compare-and-swaps that would retry at its root.  All lanes of the
// wave call this (ballot and shuffle).
__device__ __forceinline__ void face_union(int32_t *par, bool act, int32_t a, int32_t b) {
  const int lane = (int)(threadIdx.x % kWave);
  const unsigned long long m = __ballot(act);
  const int32_t pa = act ? load_par(par, a) : -1;
  const int32_t pb = act ? load_par(par, b) : -1;
  const unsigned long long below = m & ((1ull << lane) - 1ull);
  const int prev = below ? 63 - __clzll((long long)below) : lane;
  const int32_t qa = __shfl(pa, prev, kWave), qb = __shfl(pb, prev, kWave);
  if (act && !(below && qa == pa && qb == pb)) unite(par, pa, pb);
}

// Unions across tile faces, on the tile roots k_ccl_local left (a kernel
// boundary after it: every parent is written before any is hooked).
__global__ __launch_bounds__(kTile) void k_ccl_border(int32_t *par,
                                                       const uint8_t *__restrict__ mask,
                                                       int nx, int ny, int nz) {
  const TileVoxel tv = tile_voxel(nx, ny, nz);
  const bool fg = tv.in && mask[tv.v];
  const int32_t vi = (int32_t)tv.v;
  const int64_t sx = (int64_t)ny * nz;
  face_union(par, fg && tv.lz == 0 && tv.z > 0 && mask[tv.v - 1], vi, vi - 1);
  face_union(par, fg && tv.ly == 0 && tv.y > 0 && mask[tv.v - nz], vi, vi - nz);
  face_union(par, fg && tv.lx == 0 && tv.x > 0 && mask[tv.v - sx], vi, (int32_t)(tv.v - sx));
}

// After all unions (a kernel boundary): parent := root.  Every store here
// writes a root, so concurrent walks through v stay correct.
__global__ __launch_bounds__(kBlock) void k_ccl_flatten(int32_t *par, int64_t V) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= V) return;
  const int32_t p = load_par(par, (int32_t)v);
  if (p >= 0 && p != (int32_t)v)
    __hip_atomic_store(par + v, root_of(par, p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Roots per chunk of kScanChunk voxels.
__global__ __launch_bounds__(kBlock) void k_ccl_count(int32_t *__restrict__ bsum,
                                                      const int32_t *__restrict__ par,
                                                      int64_t V) {
  __shared__ int32_t sh[1][kBlock];
  const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
  int32_t n[1] = {0};
  for (int q = 0; q < kScanChunk / kBlock; ++q) {
    const int64_t v = c0 + (int64_t)q * kBlock + threadIdx.x;
    if (v < V && par[v] == (int32_t)v) ++n[0];
  }
  block_sum<int32_t, 1>(n, sh);
  if (threadIdx.x == 0) bsum[blockIdx.x] = n[0];
}

// One block: exclusive scan of the chunk counts in place, total to *count.
__global__ __launch_bounds__(kBlock) void k_ccl_scan(int32_t *__restrict__ bsum,
                                                     int32_t *__restrict__ count,
                                                     int64_t nchunks) {
  __shared__ int32_t sh[kBlock];
  int32_t carry = 0;
  for (int64_t base = 0; base < nchunks; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const int32_t x = i < nchunks ? bsum[i] : 0;
    sh[threadIdx.x] = x;
    __syncthreads();
    // Hillis-Steele inclusive scan of the tile
    for (int s = 1; s < kBlock; s <<= 1) {
      const int32_t y = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
      __syncthreads();
      sh[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < nchunks) bsum[i] = carry + sh[threadIdx.x] - x;
    carry += sh[kBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = carry;
}

// Roots get their number 1..n in index order; the rest are written by
// k_ccl_label after this kernel has finished.  The flags are 0 or 1, so a
// lane's rank within its wave is a popcount of the wave's ballot below it,
// and the waves' totals go through LDS: two barriers per 256 voxels.
__global__ __launch_bounds__(kBlock) void k_ccl_number(int32_t *__restrict__ labels,
                                                       const int32_t *__restrict__ par,
                                                       const int32_t *__restrict__ boff,
                                                       int64_t V) {
  __shared__ int32_t wsum[kBlock / kWave];
  const int lane = (int)(threadIdx.x % kWave), w = (int)(threadIdx.x / kWave);
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
  const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
  int32_t carry = boff[blockIdx.x];
  for (int q = 0; q < kScanChunk / kBlock; ++q) {
    const int64_t v = c0 + (int64_t)q * kBlock + threadIdx.x;
    const bool f = v < V && par[v] == (int32_t)v;
    const unsigned long long m = __ballot(f);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int32_t before = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; ++i) {
      if (i < w) before += wsum[i];
      total += wsum[i];
    }
    // inclusive count = 1-based number
    if (f) labels[v] = carry + before + __popcll(m & upto);
    carry += total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_ccl_label(int32_t *__restrict__ labels,
                                                      const int32_t *__restrict__ par,
                                                      int64_t V) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= V) return;
  const int32_t p = par[v];
  if (p < 0) labels[v] = 0;
  else if (p != (int32_t)v) labels[v] = labels[p];
}

// sizes[l] += 1 for every entry equal to l, l = 1..n_labels, with one
// atomic per distinct label of a wave: the wave repeatedly takes the label
// of its first pending lane and adds the count of lanes holding it.  Label 0
// (and anything outside 1..n) is left to k_sizes_rest: counting the
// background with atomics would put most voxels on one counter.
// Per-label voxel counts.  A 1024-thread block takes 4096 voxels; each wave
// counts the lanes holding one label at a time (ballot), adds the count to
// the block's LDS table (labels 1..kSizeTable) and the block adds each
// non-zero entry to sizes[] once -- a few hundred global atomics per label
// instead of one per wave (a large component's counter was the bottleneck:
// 66 us for one-atomic-per-wave on a 1.5M-voxel field).  Labels beyond the
// table go straight to sizes[] per wave; the background is not counted here.
constexpr int kSizeThreads = 1024, kSizePer = 4, kSizeTable = 4096;
__global__ __launch_bounds__(kSizeThreads) void k_label_sizes(int64_t *__restrict__ sizes,
                                                              const int32_t *__restrict__ labels,
                                                              int64_t V, int n_labels) {
  __shared__ int32_t tab[kSizeTable];
  const int lim = min(n_labels, kSizeTable);
  for (int i = threadIdx.x; i < lim; i += kSizeThreads) tab[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int64_t base = (int64_t)blockIdx.x * (kSizeThreads * kSizePer);
#pragma unroll
  for (int k = 0; k < kSizePer; ++k) {
    const int64_t v = base + (int64_t)k * kSizeThreads + threadIdx.x;
    const int32_t l = v < V ? labels[v] : 0;
    bool pending = l >= 1 && l <= n_labels;
    uint64_t pm = __ballot(pending);
    while (pm) {  // wave-uniform: at most one pass per distinct label
      const int first = __ffsll((long long)pm) - 1;
      const int32_t lf = __shfl(l, first, kWave);
      const bool mine = pending && l == lf;
      const uint64_t m = __ballot(mine);
      if (lane == first) {
        if (lf <= kSizeTable)
          atomicAdd(&tab[lf - 1], (int32_t)__popcll(m));
        else
          atomicAdd((unsigned long long *)&sizes[lf], (unsigned long long)__popcll(m));
      }
      if (mine) pending = false;
      pm = __ballot(pending);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < lim; i += kSizeThreads)
    if (tab[i]) atomicAdd((unsigned long long *)&sizes[i + 1], (unsigned long long)tab[i]);
}

// One block: sizes[0] = V - sum of sizes[1..n] (entries outside 1..n).
__global__ __launch_bounds__(kBlock) void k_sizes_rest(int64_t *__restrict__ sizes,
                                                       int64_t V, int n_labels) {
  __shared__ int64_t sh[1][kBlock];
  int64_t t[1] = {0};
  for (int l = 1 + (int)threadIdx.x; l <= n_labels; l += kBlock) t[0] += sizes[l];
  block_sum<int64_t, 1>(t, sh);
  if (threadIdx.x == 0) sizes[0] = V - t[0];
}

}  // namespace

// ---------------------------------------------------------------- C ABI --

extern "C" size_t mmn_li_threshold_workspace_size(int64_t n) {
  if (n <= 0) return 0;
  return li_layout(n, nullptr, nullptr);
}

extern "C" int mmn_li_threshold(double *threshold, int *iterations, int *converged,
                                uint8_t *below, const float *img, int64_t n,
                                double exponent, int max_iter, void *workspace,
                                size_t workspace_bytes, void *stream) {
  MMN_REQUIRE(threshold && img, MMN_ERR_NULL, "mmn_li_threshold: NULL pointer argument");
  MMN_REQUIRE(n > 0, MMN_ERR_ARG, "mmn_li_threshold: n must be positive");
  MMN_REQUIRE(max_iter > 0, MMN_ERR_ARG, "mmn_li_threshold: max_iter must be positive");
  MMN_REQUIRE(exponent > 0.0, MMN_ERR_ARG, "mmn_li_threshold: exponent must be positive");
  const size_t need = li_layout(n, nullptr, nullptr);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_li_threshold: workspace of %zu bytes needed", need);
  LiWs w;
  li_layout(n, &w, (char *)workspace);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_li_prep, dim3(kGrid), dim3(kBlock), 0, st, w.a, w.part,
                     (const float2 *)img, n, exponent);
  hipLaunchKernelGGL(k_li_start, dim3(1), dim3(kBlock), 0, st, w.st, w.part, n, max_iter);
  int rc = mmn_check_launch("mmn_li_threshold");
  if (rc) return rc;
  LiState h;
  while (true) {
    for (int k = 0; k < kItersPerBatch; ++k) {
      hipLaunchKernelGGL(k_li_sums, dim3(kGrid), dim3(kBlock), 0, st, w.part, w.a, w.st, n);
      hipLaunchKernelGGL(k_li_update, dim3(1), dim3(kBlock), 0, st, w.st, w.part);
    }
    rc = mmn_check_launch("mmn_li_threshold");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(&h, w.st, sizeof(LiState), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_li_threshold: %s", hipGetErrorString(e));
      return (int)e;
    }
    if (h.done) break;
  }
  *threshold = h.t_next + h.lo;
  if (iterations) *iterations = h.iters;
  if (converged) *converged = h.iters < h.max_iter || !(fabs(h.t_next - h.t_curr) > 0.0);
  if (below) {
    hipLaunchKernelGGL(k_li_mask, dim3(mmn_blocks(n, kBlock)), dim3(kBlock), 0, st, below,
                       w.a, n, *threshold);
    return mmn_check_launch("mmn_li_threshold");
  }
  return 0;
}

extern "C" size_t mmn_ccl6_workspace_size(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
  return ccl_layout((int64_t)nx * ny * nz, nullptr, nullptr);
}

extern "C" int mmn_ccl6(int32_t *labels, int32_t *n_components, const uint8_t *mask,
                        int nx, int ny, int nz, void *workspace, size_t workspace_bytes,
                        void *stream) {
  MMN_REQUIRE(labels && mask, MMN_ERR_NULL, "mmn_ccl6: NULL pointer argument");
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0, MMN_ERR_ARG, "mmn_ccl6: sizes must be positive");
  const int64_t V = (int64_t)nx * ny * nz;
  MMN_REQUIRE(V <= INT32_MAX, MMN_ERR_ARG, "mmn_ccl6: more than 2^31 - 1 voxels");
  const size_t need = ccl_layout(V, nullptr, nullptr);
  MMN_REQUIRE(workspace && workspace_bytes >= need, MMN_ERR_WS,
              "mmn_ccl6: workspace of %zu bytes needed", need);
  CclWs w;
  ccl_layout(V, &w, (char *)workspace);
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(mmn_blocks(V, kBlock)), b(kBlock);
  const int64_t nch = ccl_chunks(V);
  const int64_t ntiles =
      (int64_t)((nx + kTX - 1) / kTX) * ((ny + kTY - 1) / kTY) * ((nz + kTZ - 1) / kTZ);
  hipLaunchKernelGGL(k_ccl_local, dim3((unsigned)ntiles), dim3(kTile), 0, st, w.par, mask, nx,
                     ny, nz);
  hipLaunchKernelGGL(k_ccl_border, dim3((unsigned)ntiles), dim3(kTile), 0, st, w.par, mask, nx,
                     ny, nz);
  hipLaunchKernelGGL(k_ccl_flatten, g, b, 0, st, w.par, V);
  hipLaunchKernelGGL(k_ccl_count, dim3((unsigned)nch), b, 0, st, w.bsum, w.par, V);
  hipLaunchKernelGGL(k_ccl_scan, dim3(1), b, 0, st, w.bsum, w.count, nch);
  hipLaunchKernelGGL(k_ccl_number, dim3((unsigned)nch), b, 0, st, labels, w.par, w.bsum, V);
  hipLaunchKernelGGL(k_ccl_label, g, b, 0, st, labels, w.par, V);
  int rc = mmn_check_launch("mmn_ccl6");
  if (rc || !n_components) return rc;
  hipError_t e = hipMemcpyAsync(n_components, w.count, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    mmn_set_error("mmn_ccl6: %s", hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

extern "C" int mmn_label_sizes(int64_t *sizes, const int32_t *labels, int64_t n_vox,
                               int n_labels, void *stream) {
  MMN_REQUIRE(sizes && labels, MMN_ERR_NULL, "mmn_label_sizes: NULL pointer argument");
  MMN_REQUIRE(n_vox >= 0 && n_labels >= 0, MMN_ERR_ARG,
              "mmn_label_sizes: negative size");
  hipStream_t st = (hipStream_t)stream;
  const int rc = mmn_zero_async(sizes, sizeof(int64_t) * ((size_t)n_labels + 1), st,
                                "mmn_label_sizes");
  if (rc != 0) return rc;
  if (n_vox > 0)
    hipLaunchKernelGGL(k_label_sizes, dim3(mmn_blocks(n_vox, kSizeThreads * kSizePer)),
                       dim3(kSizeThreads), 0, st, sizes, labels, n_vox, n_labels);
  hipLaunchKernelGGL(k_sizes_rest, dim3(1), dim3(kBlock), 0, st, sizes, n_vox, n_labels);
  return mmn_check_launch("mmn_label_sizes");
}
