// RSDF breadth-first search, FIFO-exact, level-synchronous on the GPU.
//
// The reference (sdf.py:180-222) runs one FIFO BFS per component.  FIFO order
// is level order, and inside level L+1 it is the order in which level-L
// voxels discovered their children: a child belongs to the first frontier
// position p (then the first direction d) that reaches it.  Frontier entry p
// of a level whose first entry has queue index q0 (the number of voxels the
// reference's FIFO queue held before it) offers the key 6 (q0 + p) + d, so
// keys grow along the queue.  For every level:
//   claim  each entry offers its keys to its enterable neighbours (in the
//          grid, same label, normal not NaN: a static 6-bit mask per voxel,
//          computed once, sdf.py:169-177); a neighbour keeps the minimum
//          (atomicMin on unique keys: the winner does not depend on
//          scheduling).  A voxel reached at an earlier level already holds a
//          smaller key (a seed holds -1), so offering to it changes nothing:
//          no visited test is needed,
//   count  each entry finds which neighbours hold its keys (its won mask),
//   scan   exclusive sums of the won counts give output offsets and the next
//          frontier size,
//   emit   each entry writes its children, in direction order, at its
//          offset: the next frontier comes out sorted by key, i.e. in FIFO
//          order, and every child's rsdf uses its winning parent:
//              rsdf[c] = rsdf[p] + dot(dir * dx, normals[c])  (sdf.py:222).
// Components never share voxels (labels differ), so all components run in
// one frontier, seeded in the reference's component order; each keeps its
// own FIFO order because keys only compare voxels of the same component.
//
// Two forms of a level.  Frontiers of up to about a thousand entries run in
// k_bfs_small, a single 1024-thread workgroup that loops over levels with
// the frontier, its rsdf and its masks in LDS: no launches, and the
// frontier's data never goes through global memory, but one CU issues all
// of a level's scattered requests.  Larger frontiers run as one kernel per
// level over a grid of blocks, each block owning one contiguous chunk of the
// frontier in scan order: k_level counts each entry's won children, finds
// its output offset by a single-pass scan across blocks (decoupled
// look-back), emits, and has each child offer the next level's keys at
// once.  (The earlier form, two kernels per level -- k_count, then
// k_emit_claim summing the counts before it -- stays selectable for
// measurement.)  The host enqueues those levels in batches of 64, reads the
// frontier size back once per batch, and hands back to k_bfs_small when the
// frontier shrinks.
#include <stdlib.h>

#include <algorithm>
#include <mutex>

#include "mmn_internal.h"

namespace {

constexpr int kGrid = 1024;      // blocks of the grid-stride kernels; most level blocks
constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kSmallThreads = 1024;  // one workgroup; PER positions per thread

struct Ws {
  int32_t *claim;   // [V]  smallest key offered (-1 at seeds, INT32_MAX never)
  int32_t *front0;  // [V]
  int32_t *front1;  // [V]
  uint8_t *pmask;   // [V]  won mask of frontier position p (bit d)
  uint8_t *emask;   // [V]  enterable neighbours of each voxel (bit d)
  int32_t *bcnt;    // [kGrid] per-block counts of a level
  uint32_t *scan;   // [2][kGrid] k_level's look-back status words, by level parity
  uint32_t *ctr;    // [2] k_level's block-id counters, by level parity
  int32_t *nf;      // [8]  [0,1] frontier sizes by level parity; [2] the level
                    //      k_bfs_small stopped at; [3] non-zero when a seed
                    //      lies outside the grid; [4,5] queue index of a
                    //      level's first entry, by level parity; [6]
                    //      non-zero when a look-back wait ran out
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
  char *em = take(V);
  char *bc = take(sizeof(int32_t) * kGrid);
  char *sc = take(sizeof(uint32_t) * 2 * kGrid);
  char *ct = take(sizeof(uint32_t) * 2);
  char *nf = take(sizeof(int32_t) * 8);
  if (w) {
    w->claim = (int32_t *)c;
    w->front0 = (int32_t *)f0;
    w->front1 = (int32_t *)f1;
    w->pmask = (uint8_t *)pm;
    w->emask = (uint8_t *)em;
    w->bcnt = (int32_t *)bc;
    w->scan = (uint32_t *)sc;
    w->ctr = (uint32_t *)ct;
    w->nf = (int32_t *)nf;
  }
  return off;
}

// Size of the frontier of parity cur; 0 once k_seed has flagged a seed
// outside the grid or a k_level wait has run out, so that no level kernel
// indexes with it.
__device__ __forceinline__ int front_size(const int32_t *__restrict__ nf, int cur) {
  return (nf[3] | nf[6]) ? 0 : nf[cur];
}

__device__ __forceinline__ void chunk_of(int n, int &start, int &end) {
  const int chunk = (n + (int)gridDim.x - 1) / (int)gridDim.x;
  start = min(n, (int)blockIdx.x * chunk);
  end = min(n, start + chunk);
}

// Step of direction d in linear index.  The reference's neighbour order
// (dX, dY, dZ of sdf.py:145-147): d = 0..5 is (-1, 0, 0), (0, 1, 0), (1, 0, 0), (0, -1, 0),
// (0, 0, 1), (0, 0, -1) in (x, y, z), index x * ny * nz + y * nz + z.
__device__ __forceinline__ int32_t dir_step(int d, int ny, int nz) {
  const int32_t yz = ny * nz;
  return d == 0 ? -yz : d == 1 ? nz : d == 2 ? yz : d == 3 ? -nz : d == 4 ? 1 : -1;
}

// Bit d set when voxel u has a neighbour in direction d inside the grid.
__device__ __forceinline__ unsigned inside_mask(int32_t u, int nx, int ny, int nz) {
  const uint32_t yz = (uint32_t)(ny * nz);
  const uint32_t x = (uint32_t)u / yz, r = (uint32_t)u - x * yz;
  const uint32_t y = r / (uint32_t)nz, z = r - y * (uint32_t)nz;
  return (unsigned)(x > 0) | (unsigned)(y + 1 < (uint32_t)ny) << 1 |
         (unsigned)(x + 1 < (uint32_t)nx) << 2 | (unsigned)(y > 0) << 3 |
         (unsigned)(z + 1 < (uint32_t)nz) << 4 | (unsigned)(z > 0) << 5;
}

// a[i] with a 32-bit byte offset (every voxel index: V <= 2^31 / 6), so the
// access takes the scalar-base form with one VGPR of address.
template <typename T>
__device__ __forceinline__ T &at32(T *a, int32_t i) {
  return *(T *)((char *)a + (uint32_t)i * (uint32_t)sizeof(T));
}

// Key of frontier position p in direction d at queue index q0 + p.  Unsigned
// arithmetic: positions past the frontier may compute keys that are never used.
__device__ __forceinline__ int32_t key_of(int32_t q0, int p, int d) {
  return (int32_t)(6u * (uint32_t)(q0 + p) + (uint32_t)d);
}

// dot(dir_d * dx, n) as the reference forms it: three products, summed.
__device__ __forceinline__ double step_dot(int d, double dx, double n0, double n1, double n2) {
  const double ex = d == 0 ? -1.0 : d == 2 ? 1.0 : 0.0;
  const double ey = d == 1 ? 1.0 : d == 3 ? -1.0 : 0.0;
  const double ez = d == 4 ? 1.0 : d == 5 ? -1.0 : 0.0;
  const double t0 = __dmul_rn(ex * dx, n0);
  const double t1 = __dmul_rn(ey * dx, n1);
  const double t2 = __dmul_rn(ez * dx, n2);
  return __dadd_rn(__dadd_rn(t0, t1), t2);
}

__global__ void k_init(double *__restrict__ rsdf, int32_t *__restrict__ hops,
                       int32_t *__restrict__ claim, int32_t *__restrict__ nf, int64_t V) {
  const double qnan = __longlong_as_double(0x7ff8000000000000LL);
  if (blockIdx.x == 0 && threadIdx.x == 0) nf[3] = nf[6] = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
       i += (int64_t)gridDim.x * blockDim.x) {
    rsdf[i] = qnan;
    hops[i] = -1;
    claim[i] = INT32_MAX;
  }
}

// emask[v] bit d: neighbour d of v is in the grid, has a non-NaN normal and
// v's label (the enterability test of sdf.py:169-177 without "not visited").
__global__ void k_emask(uint8_t *__restrict__ emask, const double *__restrict__ nrm,
                        const int32_t *__restrict__ lab, int nx, int ny, int nz) {
  const int32_t V = nx * ny * nz;
  for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const unsigned in = inside_mask(v, nx, ny, nz);
    const int32_t lv = at32(lab, v);
    unsigned m = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      if (!(in >> d & 1)) continue;
      const int32_t c = v + dir_step(d, ny, nz);
      if (!isnan(nrm[3 * (int64_t)c]) && at32(lab, c) == lv) m |= 1u << d;
    }
    at32(emask, v) = (uint8_t)m;
  }
}

__global__ void k_seed(double *__restrict__ rsdf, int32_t *__restrict__ hops,
                       int32_t *__restrict__ claim, int32_t *__restrict__ front,
                       int32_t *__restrict__ nf, const int32_t *__restrict__ seeds, int n,
                       int32_t V) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const int32_t s = seeds[t];
    front[t] = s;
    if (s >= 0 && s < V) {
      hops[s] = 0;
      rsdf[s] = 0.0;
      claim[s] = -1;
    } else {
      atomicOr(&nf[3], 1);  // the host reports MMN_ERR_INPUT
    }
  }
  if (t == 0) {
    nf[0] = n;
    nf[1] = 0;
    nf[2] = 0;
    nf[4] = 0;
    nf[5] = 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_claim(
    const int32_t *__restrict__ front, const int32_t *__restrict__ nf, int cur,
    int32_t *__restrict__ claim, const uint8_t *__restrict__ emask, int ny, int nz,
    uint32_t *__restrict__ scan, uint32_t *__restrict__ ctr) {
  if (blockIdx.x == 0) {  // a clean scan state for the k_level launches that follow
    for (int b = threadIdx.x; b < 2 * kGrid; b += kBlock) scan[b] = 0u;
    if (threadIdx.x < 2) ctr[threadIdx.x] = 0u;
  }
  int start, end;
  chunk_of(front_size(nf, cur), start, end);
  const int32_t q0 = nf[4 + cur];
  for (int p = start + threadIdx.x; p < end; p += kBlock) {
    const int32_t u = front[p];
    const unsigned m = at32(emask, u);
#pragma unroll
    for (int d = 0; d < 6; ++d)
      if (m >> d & 1) atomicMin(&at32(claim, u + dir_step(d, ny, nz)), key_of(q0, p, d));
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
    const int32_t *__restrict__ front, const int32_t *__restrict__ nf, int cur,
    const int32_t *__restrict__ claim, const uint8_t *__restrict__ emask,
    uint8_t *__restrict__ pmask, int32_t *__restrict__ bcnt, int ny, int nz) {
  __shared__ int sh[kBlock / kWave];
  int start, end;
  chunk_of(front_size(nf, cur), start, end);
  const int32_t q0 = nf[4 + cur];
  int cnt = 0;
  for (int p = start + threadIdx.x; p < end; p += kBlock) {
    const int32_t u = front[p];
    const unsigned m = at32(emask, u);
    unsigned won = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d)
      if ((m >> d & 1) && at32(claim, u + dir_step(d, ny, nz)) == key_of(q0, p, d))
        won |= 1u << d;
    pmask[p] = (uint8_t)won;
    cnt += __popc(won);
  }
  const int tot = block_sum(cnt, sh);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// Emit, and offer each child's keys of the next level at once.  Each block
// finds its output offset by summing the counts of the blocks before it
// (k_count wrote them; 1-4 KB), and block 0 writes the next level's size and
// queue index into the other parity's slots (this launch reads only its
// own).  The child at next-frontier position o has queue index q1 + o,
// q1 = q0 + n.  Other blocks may still be emitting this level; a voxel they
// reach holds a smaller key, so the early offers change nothing there.
// (A "last block scans" variant of k_count needs a device-scope fence in
// every block, which on this part made a level twice as slow.)
__global__ __launch_bounds__(kBlock) void k_emit_claim(
    const int32_t *__restrict__ front, int32_t *__restrict__ nf, int cur,
    const uint8_t *__restrict__ pmask, const int32_t *__restrict__ bcnt,
    int32_t *__restrict__ next, int32_t *__restrict__ hops,
    double *__restrict__ rsdf, const double *__restrict__ nrm,
    const uint8_t *__restrict__ emask, int32_t *__restrict__ claim, double dx,
    int level, int ny, int nz) {
  __shared__ int sh[kBlock / kWave];
  const int n = front_size(nf, cur);
  int start, end;
  chunk_of(n, start, end);
  const int32_t q1 = nf[4 + cur] + n;
  int before = 0, all = 0;
  for (int b = (int)threadIdx.x; b < (int)gridDim.x; b += kBlock) {
    const int v = bcnt[b];
    all += v;
    if (b < (int)blockIdx.x) before += v;
  }
  int base = block_sum(before, sh);
  if (blockIdx.x == 0) {
    const int total = block_sum(all, sh);
    if (threadIdx.x == 0) {
      nf[cur ^ 1] = total;
      nf[4 + (cur ^ 1)] = q1;
    }
  }
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
      const double prev = at32(rsdf, u);
      for (int d = 0; d < 6; ++d) {
        if (!(m & (1u << d))) continue;
        const int32_t ch = u + dir_step(d, ny, nz);
        const double *nc = nrm + 3 * (int64_t)ch;
        const unsigned em = at32(emask, ch);
        next[out] = ch;
        at32(hops, ch) = level + 1;
        at32(rsdf, ch) = __dadd_rn(prev, step_dot(d, dx, nc[0], nc[1], nc[2]));
#pragma unroll
        for (int e = 0; e < 6; ++e)
          if (em >> e & 1) atomicMin(&at32(claim, ch + dir_step(e, ny, nz)), key_of(q1, out, e));
        ++out;
      }
    }
    base += round_tot;
  }
}

// One level in one launch: k_count's won masks, the offsets of k_emit_claim
// from a single-pass scan across blocks (decoupled look-back), and its
// emission with the next level's claims.
//
// A block takes its position in the scan from a counter (atomicAdd at its
// start), not from blockIdx, so every block it waits for has already
// started and, waiting only for blocks before it, finishes: the waits end.
// Block id takes frontier chunk id.  After counting its children it
// publishes the count in its status word (state "aggregate"), then sums the
// words of the blocks before it, nearest first, 64 at a time (one per lane
// of wave 0), until it meets a block whose word holds an inclusive prefix;
// then it publishes its own inclusive prefix.  A status word is one 32-bit
// value, state in bits 30-31 (0 none, 1 aggregate, 2 inclusive prefix) and
// the count in bits 0-29 (a level has fewer than 2^30 children), written
// and read with device-scope atomic read-modify-writes, the operations the
// claims already rely on across XCDs; no other data crosses blocks inside
// the launch.  A wait that exceeds spin_limit reads sets nf[6] and the block
// emits nothing; the host reports it.  The last block in scan order writes
// the next level's size and queue index into the other parity's slots, and
// block 0 clears the other parity's scan state for the next launch.  Won
// masks of a block's first kLevelLds positions stay in LDS between the two
// passes; later positions are tested again, which gives the same answer:
// this level's keys are final, and the next level's offers made meanwhile
// are larger and cannot replace them.
constexpr int kLevelLds = 4096;
constexpr uint32_t kScanAgg = 1u << 30, kScanPre = 2u << 30, kScanVal = (1u << 30) - 1u;

// The directions of an entry that hold its keys.  All six claims are read
// unconditionally (a direction that is not enterable reads u's own claim,
// which is never this level's key) so that the six loads issue together
// instead of one per branch.
__device__ __forceinline__ unsigned won_mask(int32_t u, unsigned m, int32_t q0, int p,
                                             const int32_t *__restrict__ claim, int ny,
                                             int nz) {
  int32_t k[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) k[d] = at32(claim, (m >> d & 1) ? u + dir_step(d, ny, nz) : u);
  unsigned won = 0;
#pragma unroll
  for (int d = 0; d < 6; ++d)
    won |= ((m >> d & 1) && k[d] == key_of(q0, p, d)) ? 1u << d : 0u;
  return won;
}

__global__ __launch_bounds__(kBlock) void k_level(
    const int32_t *__restrict__ front, int32_t *__restrict__ nf, int cur,
    uint32_t *__restrict__ scan, uint32_t *__restrict__ ctr, int32_t *__restrict__ next,
    int32_t *__restrict__ hops, double *__restrict__ rsdf, const double *__restrict__ nrm,
    const uint8_t *__restrict__ emask, int32_t *__restrict__ claim, double dx, int level,
    int ny, int nz, int spin_limit) {
  __shared__ int sh[kBlock / kWave];
  __shared__ int s_id, s_excl;
  __shared__ uint8_t s_won[kLevelLds];
  const int nxt = cur ^ 1;
  const int G = (int)gridDim.x;
  // the frontier size first: it does not depend on the block's scan position
  const int n = front_size(nf, cur);
  const int32_t q0 = nf[4 + cur];
  if (n == 0) {  // past the end of the search (or a flagged error): nothing to scan
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      nf[nxt] = 0;
      nf[4 + nxt] = q0;
    }
    return;
  }
  if (threadIdx.x == 0) s_id = (int)atomicAdd(&ctr[cur], 1u);
  __syncthreads();
  const int id = s_id;
  if (id == 0) {
    for (int b = threadIdx.x; b < G; b += kBlock) scan[nxt * kGrid + b] = 0u;
    if (threadIdx.x == 0) ctr[nxt] = 0u;
  }
  const int chunk = (n + G - 1) / G;
  const int start = min(n, id * chunk), end = min(n, start + chunk);
  // pass 1: won masks and the block's child count.  For the first round of
  // positions the children's data (parent rsdf, child normal and mask) are
  // loaded after the count's barrier (which waits for outstanding loads)
  // and before the scan, so that they overlap it; only the stores and the
  // next level's offers wait for the scan.
  int cnt = 0;
  int32_t u0 = 0;
  unsigned won0 = 0u;
  double prev0 = 0.0, c0[6], c1[6], c2[6];
  unsigned em0[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    c0[d] = c1[d] = c2[d] = 0.0;
    em0[d] = 0u;
  }
  for (int r0 = start; r0 < end; r0 += kBlock) {
    const int p = r0 + (int)threadIdx.x;
    if (p >= end) continue;
    const int32_t u = front[p];
    const unsigned won = won_mask(u, at32(emask, u), q0, p, claim, ny, nz);
    if (p - start < kLevelLds) s_won[p - start] = (uint8_t)won;
    cnt += __popc(won);
    if (r0 == start) {
      u0 = u;
      won0 = won;
    }
  }
  const int tot = block_sum(cnt, sh);
  // unconditional loads (a direction not won reads u0's own entries): one
  // burst of requests in flight during the scan
  prev0 = at32(rsdf, u0);
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const int32_t ch = (won0 >> d & 1) ? u0 + dir_step(d, ny, nz) : u0;
    const double *nc = nrm + 3 * (int64_t)ch;
    c0[d] = nc[0];
    c1[d] = nc[1];
    c2[d] = nc[2];
    em0[d] = at32(emask, ch);
  }
  // the scan across blocks (wave 0)
  uint32_t *__restrict__ st = scan + cur * kGrid;
  if (threadIdx.x < kWave) {
    const int l = (int)threadIdx.x;
    int excl = 0;
    bool ok = true;
    if (id == 0) {
      if (l == 0) atomicExch(&st[0], kScanPre | (uint32_t)tot);
    } else {
      if (l == 0) atomicExch(&st[id], kScanAgg | (uint32_t)tot);
      int base = id - 1, spins = 0;
      for (;;) {
        const int j = base - l;
        const uint32_t wv = j >= 0 ? atomicOr(&st[j], 0u) : kScanPre;
        const uint32_t state = wv & ~kScanVal;
        const unsigned long long ready = __ballot(state != 0u);
        const unsigned long long pre = __ballot(state == kScanPre);
        // lanes up to the nearest inclusive prefix (or all 64) must be ready
        const int upto = pre ? __ffsll((long long)pre) - 1 : kWave - 1;
        const unsigned long long need =
            upto == kWave - 1 ? ~0ull : ((1ull << (upto + 1)) - 1ull);
        if ((ready & need) == need) {
          int v = l <= upto ? (int)(wv & kScanVal) : 0;
          for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
          excl += v;
          if (pre) break;
          base -= kWave;
          continue;
        }
        if (++spins > spin_limit) {
          ok = false;
          break;
        }
      }
      if (l == 0) {
        if (ok)
          atomicExch(&st[id], kScanPre | (uint32_t)(excl + tot));
        else
          atomicOr((uint32_t *)&nf[6], 1u);
      }
    }
    if (l == 0) {
      s_excl = ok ? excl : -1;
      if (ok && id == G - 1) {  // the next level's size and queue index
        nf[nxt] = excl + tot;
        nf[4 + nxt] = q0 + n;
      }
    }
  }
  __syncthreads();
  if (s_excl < 0) return;
  // pass 2: emit in position order and offer the children's next-level keys
  const int32_t q1 = q0 + n;
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  int base = s_excl;
  for (int r0 = start; r0 < end; r0 += kBlock) {
    const int p = r0 + (int)threadIdx.x;
    const bool first = r0 == start;
    unsigned m = 0u;
    int32_t u = 0;
    if (first) {
      m = won0;
      u = u0;
    } else if (p < end) {
      u = front[p];
      m = p - start < kLevelLds ? (unsigned)s_won[p - start]
                                : won_mask(u, at32(emask, u), q0, p, claim, ny, nz);
    }
    const int c = __popc(m);
    int inc = c;
    for (int o = 1; o < kWave; o <<= 1) {
      const int t = __shfl_up(inc, o, kWave);
      if (l >= o) inc += t;
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
      if (first) {
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          if (!(m >> d & 1)) continue;
          const int32_t ch = u + dir_step(d, ny, nz);
          next[out] = ch;
          at32(hops, ch) = level + 1;
          at32(rsdf, ch) = __dadd_rn(prev0, step_dot(d, dx, c0[d], c1[d], c2[d]));
#pragma unroll
          for (int e = 0; e < 6; ++e)
            if (em0[d] >> e & 1)
              atomicMin(&at32(claim, ch + dir_step(e, ny, nz)), key_of(q1, out, e));
          ++out;
        }
      } else {
        const double prev = at32(rsdf, u);
        for (int d = 0; d < 6; ++d) {
          if (!(m & (1u << d))) continue;
          const int32_t ch = u + dir_step(d, ny, nz);
          const double *nc = nrm + 3 * (int64_t)ch;
          const unsigned em = at32(emask, ch);
          next[out] = ch;
          at32(hops, ch) = level + 1;
          at32(rsdf, ch) = __dadd_rn(prev, step_dot(d, dx, nc[0], nc[1], nc[2]));
#pragma unroll
          for (int e = 0; e < 6; ++e)
            if (em >> e & 1)
              atomicMin(&at32(claim, ch + dir_step(e, ny, nz)), key_of(q1, out, e));
          ++out;
        }
      }
    }
    base += round_tot;
  }
}

// Levels level0, level0 + 1, ... in one workgroup while the frontier has
// 1..cap entries (cap <= 1024 PER).  Writes each level's next size to
// nf[parity], and at exit the first level it did not run to nf[2] and that
// level's queue index to nf[4 + parity].  The frontier (voxel, rsdf, enterable mask)
// stays in LDS between levels; thread t owns positions [PER t, PER t + PER)
// in the claim, count and scan phases (so an exclusive scan over threads
// keeps position order) and entries t, t + 1024, ... when the next frontier
// is completed.  Positions past the frontier reread its last entry and are
// masked off, so every load of a phase is issued before any is used.  claim[]
// is read with agent-scope atomic loads (the atomics are performed at L2; a
// plain load could hit an older L1 line).  hops and rsdf are written here
// and read only by later launches.  A level whose next frontier outgrows the
// workgroup writes it to global memory for the multi-block kernels and ends
// the loop.
template <int PER>
__global__ __launch_bounds__(kSmallThreads) void k_bfs_small(
    int32_t *__restrict__ fr0, int32_t *__restrict__ fr1, int32_t *__restrict__ nf,
    int32_t *__restrict__ claim, int32_t *__restrict__ hops, double *__restrict__ rsdf,
    const double *__restrict__ nrm, const uint8_t *__restrict__ emask, double dx, int level0,
    int cap, int ny, int nz) {
  constexpr int kCap = kSmallThreads * PER;
  __shared__ int32_t s_u[kCap];  // frontier voxels
  __shared__ double s_r[kCap];   // their rsdf (parent rsdf while a level is emitted)
  __shared__ uint8_t s_m[kCap];  // their enterable masks
  __shared__ uint8_t s_d[kCap];  // direction each next-frontier entry was reached in
  __shared__ int wsum[kSmallThreads / kWave];
  const int t = threadIdx.x;
  const int w = t / kWave, l = t % kWave;
  const int p0 = t * PER;
  int level = level0;
  int n = front_size(nf, level & 1);
  int32_t q0 = nf[4 + (level & 1)];
  if (n > 0 && n <= cap) {
    const int32_t *__restrict__ cur = (level & 1) ? fr1 : fr0;
    for (int j = t; j < n; j += kSmallThreads) {
      const int32_t u = cur[j];
      s_u[j] = u;
      s_r[j] = at32(rsdf, u);
      s_m[j] = at32(emask, u);
    }
  }
  __syncthreads();
  while (n > 0 && n <= cap) {
    int32_t *__restrict__ nxt = (level & 1) ? fr0 : fr1;
    int32_t u[PER];
    unsigned m[PER];
    double pr[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int pc = min(p0 + q, n - 1);
      u[q] = s_u[pc];
      m[q] = p0 + q < n ? (unsigned)s_m[pc] : 0u;
      pr[q] = s_r[pc];
    }
    // claim
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int d = 0; d < 6; ++d)
        if (m[q] >> d & 1)
          atomicMin(&at32(claim, u[q] + dir_step(d, ny, nz)), key_of(q0, p0 + q, d));
    __syncthreads();
    // count
    unsigned won[PER];
    int cnt = 0;
    // one request per enterable neighbour (a single CU's memory pipeline
    // bounds a level), all issued before any is compared: the empty asm
    // takes the six values, so the compiler cannot fold a compare into a
    // load's branch and wait for that load before issuing the next
    int32_t k[PER][6];
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int d = 0; d < 6; ++d)
        k[q][d] = (m[q] >> d & 1)
                      ? __hip_atomic_load(&at32(claim, u[q] + dir_step(d, ny, nz)),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      : -1;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      asm volatile("" : "+v"(k[q][0]), "+v"(k[q][1]), "+v"(k[q][2]), "+v"(k[q][3]),
                   "+v"(k[q][4]), "+v"(k[q][5]));
      unsigned wm = 0;
#pragma unroll
      for (int d = 0; d < 6; ++d)
        if ((m[q] >> d & 1) && k[q][d] == key_of(q0, p0 + q, d)) wm |= 1u << d;
      won[q] = wm;
      cnt += __popc(wm);
    }
    // exclusive scan of the threads' counts (thread order = position order)
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const int v = __shfl_up(inc, o, kWave);
      if (l >= o) inc += v;
    }
    if (l == kWave - 1) wsum[w] = inc;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kSmallThreads / kWave; ++i) {
      if (i < w) before += wsum[i];
      total += wsum[i];
    }
    int out = before + inc - cnt;
    if (total <= cap) {
      // emit: the children, in key order, with their parents' rsdf ...
#pragma unroll
      for (int q = 0; q < PER; ++q)
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          if (!(won[q] >> d & 1)) continue;
          const int32_t ch = u[q] + dir_step(d, ny, nz);
          s_u[out] = ch;
          s_r[out] = pr[q];
          s_d[out] = (uint8_t)d;
          at32(nxt, out) = ch;
          ++out;
        }
      __syncthreads();
      // ... then each child's normal, rsdf and mask
      if (total > 0) {
        int32_t ch[PER];
        int dd[PER];
        double n0[PER], n1[PER], n2[PER];
        unsigned em[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
          const int j = min(t + k * kSmallThreads, total - 1);
          ch[k] = s_u[j];
          dd[k] = s_d[j];
          const double *nc = nrm + 3 * (int64_t)ch[k];
          n0[k] = nc[0];
          n1[k] = nc[1];
          n2[k] = nc[2];
          em[k] = at32(emask, ch[k]);
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
          const int j = t + k * kSmallThreads;
          if (j >= total) continue;
          const double r = __dadd_rn(s_r[j], step_dot(dd[k], dx, n0[k], n1[k], n2[k]));
          s_r[j] = r;
          s_m[j] = (uint8_t)em[k];
          at32(hops, ch[k]) = level + 1;
          at32(rsdf, ch[k]) = r;
        }
      }
    } else {
      // the next frontier outgrows the workgroup: write it for the multi-block kernels
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (!won[q]) continue;
        for (int d = 0; d < 6; ++d) {
          if (!(won[q] >> d & 1)) continue;
          const int32_t ch = u[q] + dir_step(d, ny, nz);
          const double *nc = nrm + 3 * (int64_t)ch;
          at32(nxt, out++) = ch;
          at32(hops, ch) = level + 1;
          at32(rsdf, ch) = __dadd_rn(pr[q], step_dot(d, dx, nc[0], nc[1], nc[2]));
        }
      }
    }
    if (t == 0) nf[(level + 1) & 1] = total;
    q0 += n;
    ++level;
    n = total;
    // the next level rereads s_u, s_r, s_m written above and rewrites wsum
    __syncthreads();
  }
  if (t == 0) {
    nf[2] = level;
    nf[4 + (level & 1)] = q0;
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
                     w.claim, w.nf, V);
  hipLaunchKernelGGL(k_seed, dim3(mmn_blocks(n_seeds > 0 ? n_seeds : 1, kBlock)),
                     dim3(kBlock), 0, st, rsdf, hops, w.claim, w.front0, w.nf, seeds,
                     n_seeds, (int32_t)V);
  int rc = mmn_check_launch("mmn_rsdf_bfs(init)");
  if (rc) return rc;
  if (n_seeds == 0) return 0;
  hipLaunchKernelGGL(k_emask, dim3(kGrid), dim3(kBlock), 0, st, w.emask, normals, labels,
                     nx, ny, nz);

  // Positions per thread of the single-workgroup loop (MMN_BFS_PER: 1, 2, 4,
  // 8, or 0 for the multi-block kernels at every level).  One position per
  // thread spreads a level's scattered requests over all 16 waves: on fields
  // of 57,483 / 392,019 reached voxels (64 / 307 levels) PER 1 takes 0.66 /
  // 3.3-3.4 ms, PER 2 0.71 / 3.45-3.6 ms.
  const int per = [] {
    const char *e = getenv("MMN_BFS_PER");
    const int v = e ? atoi(e) : 1;
    return v == 0 || v == 2 || v == 4 || v == 8 ? v : 1;
  }();
  // Largest frontier the single workgroup takes (MMN_BFS_CAP, default 1024,
  // at most 1024 * per).  Its levels are bound by one CU's memory pipeline
  // (about 1.5 ns per scattered request, some 14 requests per entry), the
  // multi-block kernels' by two launches (about 10 us a level): on a
  // 1.5M-voxel field caps of 512 and 1024 give the same time.
  const int cap_env = [] {
    const char *e = getenv("MMN_BFS_CAP");
    return e ? atoi(e) : 1024;
  }();
  const int small_cap = std::max(1, std::min(cap_env, kSmallThreads * per));
  // Multi-block levels enqueued between two reads of the frontier size
  // (MMN_BFS_BATCH, default 64, at most 256).  A read costs about as much as
  // 6-10 empty levels (an empty k_level returns at once, ~3 us): with k_level
  // on the 1.5M-voxel field 32 / 64 / 128 levels per batch take 1.02 / 1.00 /
  // 1.20 ms, on the 8.4M-voxel one 2.89 / 2.92 / 2.85 ms (256: 2.67).
  const int batch = [] {
    const char *e = getenv("MMN_BFS_BATCH");
    const int v = e ? atoi(e) : 64;
    return v < 1 ? 1 : v > 256 ? 256 : v;
  }();
  // Blocks of the multi-block level kernels (MMN_BFS_GRID: 256, 512 or
  // 1024; default 256).  A level's frontier has a few thousand entries; with
  // 256 or 1024 blocks a level's kernels take the same time (4.7 + 5.7 us on
  // the 1.5M-voxel field), and 256 reads a quarter of the counts.
  const int lgrid = [] {
    const char *e = getenv("MMN_BFS_GRID");
    const int v = e ? atoi(e) : 256;
    return v == 512 || v == kGrid ? v : 256;
  }();
  // One launch per multi-block level (k_level, MMN_BFS_FUSED=1, the default)
  // or two (k_count, k_emit_claim; MMN_BFS_FUSED=0, kept for measurement).
  const bool fused = [] {
    const char *e = getenv("MMN_BFS_FUSED");
    return !(e && atoi(e) == 0);
  }();
  // Blocks of k_level (MMN_BFS_FGRID, 16-1024, default 64): the look-back
  // covers 64 blocks per round trip.
  const int fgrid = [] {
    const char *e = getenv("MMN_BFS_FGRID");
    const int v = e ? atoi(e) : 64;
    return v < 16 ? 16 : v > kGrid ? kGrid : v;
  }();
  // Reads of a status word one k_level block makes before giving up (about
  // a second); the blocks it waits for have all started, so a well-formed
  // launch never comes near it.
  constexpr int kSpinLimit = 1 << 20;
  int32_t *fr[2] = {w.front0, w.front1};
  int32_t n_cur = n_seeds;
  for (int level = 0; level <= V && n_cur > 0;) {
    if (n_cur <= small_cap) {
      // small frontiers: one workgroup runs levels until it outgrows it
      auto small = per == 8   ? k_bfs_small<8>
                   : per == 4 ? k_bfs_small<4>
                   : per == 2 ? k_bfs_small<2>
                              : k_bfs_small<1>;
      hipLaunchKernelGGL(small, dim3(1), dim3(kSmallThreads), 0, st, w.front0, w.front1, w.nf,
                         w.claim, hops, rsdf, normals, w.emask, dx, level, small_cap, ny, nz);
      rc = mmn_check_launch("mmn_rsdf_bfs(small)");
      if (rc) return rc;
      int32_t st4[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      hipError_t e = hipMemcpyAsync(st4, w.nf, sizeof(st4), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) {
        mmn_set_error("mmn_rsdf_bfs: %s", hipGetErrorString(e));
        return (int)e;
      }
      MMN_REQUIRE(st4[3] == 0, MMN_ERR_INPUT,
                  "mmn_rsdf_bfs: a seed lies outside the grid of %lld voxels", (long long)V);
      if (st4[6]) {
        mmn_set_error("mmn_rsdf_bfs: a level's scan wait ran out");
        return (int)hipErrorLaunchTimeOut;
      }
      level = st4[2];
      n_cur = st4[level & 1];
      continue;
    }
    // the batch's first level is claimed here; every emit claims the next
    // (a level that was already claimed gets the same keys again: no change)
    hipLaunchKernelGGL(k_claim, dim3(lgrid), dim3(kBlock), 0, st, fr[level & 1], w.nf,
                       level & 1, w.claim, w.emask, ny, nz, w.scan, w.ctr);
    for (int b = 0; b < batch; ++b, ++level) {
      const int cur = level & 1, nxt = cur ^ 1;
      if (fused) {
        hipLaunchKernelGGL(k_level, dim3(fgrid), dim3(kBlock), 0, st, fr[cur], w.nf, cur,
                           w.scan, w.ctr, fr[nxt], hops, rsdf, normals, w.emask, w.claim, dx,
                           level, ny, nz, kSpinLimit);
        continue;
      }
      hipLaunchKernelGGL(k_count, dim3(lgrid), dim3(kBlock), 0, st, fr[cur], w.nf, cur,
                         w.claim, w.emask, w.pmask, w.bcnt, ny, nz);
      hipLaunchKernelGGL(k_emit_claim, dim3(lgrid), dim3(kBlock), 0, st, fr[cur], w.nf, cur,
                         w.pmask, w.bcnt, fr[nxt], hops, rsdf, normals, w.emask, w.claim, dx,
                         level, ny, nz);
    }
    rc = mmn_check_launch("mmn_rsdf_bfs(level)");
    if (rc) return rc;
    int32_t st4[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    hipError_t e = hipMemcpyAsync(st4, w.nf, sizeof(st4), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      mmn_set_error("mmn_rsdf_bfs: %s", hipGetErrorString(e));
      return (int)e;
    }
    MMN_REQUIRE(st4[3] == 0, MMN_ERR_INPUT,
                "mmn_rsdf_bfs: a seed lies outside the grid of %lld voxels", (long long)V);
    if (st4[6]) {
      mmn_set_error("mmn_rsdf_bfs: a level's scan wait ran out");
      return (int)hipErrorLaunchTimeOut;
    }
    n_cur = st4[level & 1];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// No-allocation variant.  The BFS's scratch (claims, the two frontiers, the
// won and enterable masks, the per-block counts and the level counters) sits
// in one file-scope device array instead of a caller workspace.  Every GPU
// the library runs on holds its own copy of the array until the process
// ends.  In host code `scratch` names a host-side stand-in, not device
// memory: the device address comes from hipGetSymbolAddress, resolved once
// per device on the first call made with that device current, and only that
// address is passed to the kernels.  Everything the kernels read is written
// first by k_init / k_seed / k_emask / the level kernels within the same
// call, so the contents a previous call left behind never reach a result.
// ---------------------------------------------------------------------------
static constexpr size_t kBfsScratchBytes = MMN_RSDF_BFS_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kBfsScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (resolved on first use).
int scratch_on_current_device(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    mmn_set_error("mmn_rsdf_bfs_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    return (int)e;
  }
  MMN_REQUIRE(dev >= 0 && dev < kMaxDevices, MMN_ERR_ARG,
              "mmn_rsdf_bfs_noalloc: device %d beyond the %d this library tracks", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      mmn_set_error("mmn_rsdf_bfs_noalloc: hipGetSymbolAddress(scratch): %s",
                    hipGetErrorString(e));
      return e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int mmn_rsdf_bfs_noalloc(double *rsdf, int32_t *hops, const double *normals,
                                    const int32_t *labels, const int32_t *seeds,
                                    int n_seeds, double dx, int nx, int ny, int nz,
                                    void *stream) {
  MMN_REQUIRE(nx > 0 && ny > 0 && nz > 0, MMN_ERR_ARG,
              "mmn_rsdf_bfs_noalloc: sizes must be positive");
  const int64_t V = (int64_t)nx * ny * nz;
  MMN_REQUIRE(V <= (int64_t)(INT32_MAX / 6), MMN_ERR_ARG,
              "mmn_rsdf_bfs_noalloc: grid of %lld voxels exceeds the int32 key range",
              (long long)V);
  const size_t need = ws_layout(V, nullptr, nullptr);
  MMN_REQUIRE(need <= kBfsScratchBytes, MMN_ERR_ARG,
              "mmn_rsdf_bfs_noalloc: a grid of %lld voxels needs %zu bytes of scratch, more "
              "than the %zu this entry point holds (about %lld voxels); use mmn_rsdf_bfs "
              "with a workspace",
              (long long)V, need, kBfsScratchBytes, (long long)(kBfsScratchBytes / 14));
  void *ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  return mmn_rsdf_bfs(rsdf, hops, normals, labels, seeds, n_seeds, dx, nx, ny, nz, ws,
                      kBfsScratchBytes, stream);
}
