/*
 * CPU oracle for the mmNorm image-formation hot path.  TEST INFRASTRUCTURE
 * ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, as the checker and the CPU baseline -- never by the
 * product path (mmnorm_amd/), which has no CPU fallback.
 *
 * Built with -ffp-contract=off (no FMA contraction) so each product and sum
 * rounds on its own, as the reference's expressions are written.
 *
 *   or_bp_ref      restates the `cuda_image` kernel (imaging_gpu.cu:48-122):
 *                  FP32 distance, FP64 phase, FP64 cos/sin narrowed to FP32,
 *                  FP32 complex sums in row order; both branches.
 *   or_sn_ref      restates the `cuda_sn_est` kernel
 *                  (surface_normal_est.cu:64-149).
 *   or_bp_f64      the FP64 oracle O64 of the full-frequency sum: distance,
 *                  phase, cos/sin and sums all in FP64 from the FP32 inputs,
 *                  plus the vote moment sum_i (p_i + o_i/2 - v) c_i.
 *   or_bp_cpu_imaging  restates the reference's CPU imaging loop nest
 *                  (imaging.cpp:60-139, full branch, OpenMP over x slices):
 *                  the host-core baseline that bench.py times.
 *   or_bfs_fifo    restates the FIFO breadth-first search of
 *                  _construct_rsdf_components (sdf.py:180-222) for given
 *                  seeds.
 *   or_li_threshold  restates the Li iteration of skimage.filters.threshold_li
 *                  (Li & Tam's minimum cross-entropy threshold, called at
 *                  sdf.py:60-70) in FP64 with sequential sums.
 *   or_ray_matching, or_simulate_channel  restate the ray marcher
 *                  (ray_tracing_kernel.cu:60-421) and the channel simulation
 *                  of isosurface_optimization.py:111-245 (refine, FP64
 *                  distance, one FP64 exp per hit and sample).
 *   or_ccl6        6-connected component labelling (the partition of
 *                  cc3d.connected_components(mask, connectivity=6),
 *                  sdf.py:89) by flood fill in C-order scan, numbered in
 *                  order of first appearance.
 *
 * Pins (tests/test_oracle.py, tests/test_oracle_interp.py): or_bp_ref's
 * interpolated branch reads, on hand-constructed distances, the bin integer
 * arithmetic gives (at, just below and just above bin edges, with and
 * without the TI offset), skips exactly the rows with d > fs * S, carries
 * the phasor exp(-j 2 pi d / lambda), and on an FFT'd point scatterer gives
 * the geometric-sum closed form at the focus; or_sn_ref gives a single
 * row's vote (p + o/2 - v) Re(c conj I)/|I| and the normal (0, 0, 1) of a
 * symmetric four-RX aperture; both also equal their pure-Python
 * restatements bit for bit on the golden inputs; or_bp_f64
 * equals a NumPy brute force and the point-scatterer closed form;
 * or_bfs_fifo equals the deque restatement and the constant-normal closed
 * form; or_li_threshold meets its fixed-point equation and the constant and
 * two-level closed forms; or_ccl6 equals scipy.ndimage.label; the ray
 * marcher equals its pure-Python restatement and hits a plane SDF where the
 * geometry puts it (a single vertical ray exactly), and a single hit's
 * channel is exp(+j 2 pi f d / c).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define C0 2.99792458e8
static const double PI_D = 3.14159265358979323846264338327950288;

static void coords(const float *xl, const float *yl, const float *zl, int ny,
                   int nz, int64_t lin, float *vx, float *vy, float *vz) {
  int64_t yz = (int64_t)ny * nz;
  int64_t x = lin / yz, r = lin - x * yz;
  int64_t y = r / nz, z = r - y * nz;
  *vx = xl[x];
  *vy = yl[y];
  *vz = zl[z];
}

/* FP32 round-trip distance as imaging_gpu.cu:77-88 forms it. */
static float dist_ref32(float vx, float vy, float vz, const float *p,
                        const float *o, int is_ti) {
  float ax = vx - p[0], ay = vy - p[1], az = vz - p[2];
  float fwd = sqrtf(ax * ax + ay * ay + az * az);
  float bx = ax - o[0], by = ay - o[1], bz = az - o[2];
  float back = sqrtf(bx * bx + by * by + bz * bz);
  float d = fwd + back;
  if (is_ti) d = (float)((double)d + 0.15);
  return d;
}

/* imaging_gpu.cu:48-122.  meas: rows x S complex (interleaved float pairs). */
void or_bp_ref(float *img, const float *xl, const float *yl, const float *zl,
               const float *ant, const float *meas, const float *rxo,
               float slope, float wavelength, float fft_spacing, int nx,
               int ny, int nz, int rows, int R, int S, int is_ti, int interp,
               const int64_t *vidx, int64_t nsel) {
  int64_t n = vidx ? nsel : (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t t = 0; t < n; ++t) {
    int64_t lin = vidx ? vidx[t] : t;
    float vx, vy, vz;
    coords(xl, yl, zl, ny, nz, lin, &vx, &vy, &vz);
    float sr = 0.0f, si = 0.0f;
    for (int i = 0; i < rows; ++i) {
      float d = dist_ref32(vx, vy, vz, ant + 3 * i, rxo + 3 * (i % R), is_ti);
      const float *row = meas + (size_t)2 * i * S;
      if (interp) {
        if (d < 0.0f || d > fft_spacing * (float)S) continue;
        int bin = (int)floorf(d / fft_spacing / 2.0f);
        float mr = row[2 * bin], mi = row[2 * bin + 1];
        double th = -2.0 * PI_D * (double)d / (double)wavelength;
        float pr = (float)cos(th), pi = (float)sin(th);
        sr += mr * pr - mi * pi;
        si += mr * pi + mi * pr;
      } else {
        double a0 = -2.0 * PI_D * (double)d / (double)wavelength;
        for (int j = 0; j < S; ++j) {
          float mr = row[2 * j], mi = row[2 * j + 1];
          double th = a0 + -2.0 * PI_D * (double)d * (double)j * (double)slope / C0;
          float pr = (float)cos(th), pi = (float)sin(th);
          sr += mr * pr - mi * pi;
          si += mr * pi + mi * pr;
        }
      }
    }
    img[2 * t] = sr;
    img[2 * t + 1] = si;
  }
}

/* surface_normal_est.cu:64-149.  img indexed by linear voxel index. */
void or_sn_ref(float *out, const int64_t *vidx, int64_t nvalid,
               const float *img, const float *xl, const float *yl,
               const float *zl, const float *ant, const float *meas,
               const float *rxo, float wavelength, float fft_spacing, int ny,
               int nz, int rows, int R, int S, int is_ti) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t t = 0; t < nvalid; ++t) {
    int64_t lin = vidx[t];
    float vx, vy, vz;
    coords(xl, yl, zl, ny, nz, lin, &vx, &vy, &vz);
    float ir = img[2 * lin], ii = img[2 * lin + 1];
    float mag = sqrtf(ir * ir + ii * ii);
    float snx = 0.0f, sny = 0.0f, snz = 0.0f;
    for (int i = 0; i < rows; ++i) {
      const float *p = ant + 3 * i, *o = rxo + 3 * (i % R);
      float d = dist_ref32(vx, vy, vz, p, o, is_ti);
      if (d < 0.0f || d > fft_spacing * (float)S) continue;
      int bin = (int)floorf(d / fft_spacing / 2.0f);
      const float *row = meas + (size_t)2 * i * S;
      float mr = row[2 * bin], mi = row[2 * bin + 1];
      double th = -2.0 * PI_D * (double)d / (double)wavelength;
      float pr = (float)cos(th), pi = (float)sin(th);
      float cr = mr * pr - mi * pi;
      float ci = mr * pi + mi * pr;
      float vecx = (p[0] + o[0] / 2.0f) - vx;
      float vecy = (p[1] + o[1] / 2.0f) - vy;
      float vecz = (p[2] + o[2] / 2.0f) - vz;
      float w = (cr * ir) + (ci * ii);
      w /= mag;
      snx += vecx * w;
      sny += vecy * w;
      snz += vecz * w;
    }
    out[3 * t] = snx;
    out[3 * t + 1] = sny;
    out[3 * t + 2] = snz;
  }
}

/* FP64 oracle O64.  out: n complex doubles; mom (optional): n x 3 complex. */
void or_bp_f64(double *out, double *mom, const float *xl, const float *yl,
               const float *zl, const float *ant, const float *meas,
               const float *rxo, double f0, double df, int nx, int ny, int nz,
               int rows, int R, int F, int is_ti, const int64_t *vidx,
               int64_t nsel) {
  int64_t n = vidx ? nsel : (int64_t)nx * ny * nz;
  const double k0 = 2.0 * PI_D * f0 / C0, dk = 2.0 * PI_D * df / C0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < n; ++t) {
    int64_t lin = vidx ? vidx[t] : t;
    float fx, fy, fz;
    coords(xl, yl, zl, ny, nz, lin, &fx, &fy, &fz);
    double vx = fx, vy = fy, vz = fz;
    double ir = 0.0, ii = 0.0, m[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < rows; ++i) {
      const float *p = ant + 3 * i, *o = rxo + 3 * (i % R);
      double ax = vx - p[0], ay = vy - p[1], az = vz - p[2];
      double bx = ax - o[0], by = ay - o[1], bz = az - o[2];
      double d = sqrt(ax * ax + ay * ay + az * az) + sqrt(bx * bx + by * by + bz * bz);
      if (is_ti) d += 0.15;
      const float *row = meas + (size_t)2 * i * F;
      double cr = 0.0, ci = 0.0;
      for (int j = 0; j < F; ++j) {
        double th = (k0 + (double)j * dk) * d;
        double c = cos(th), s = sin(th);
        double mr = row[2 * j], mi = row[2 * j + 1];
        /* s * exp(-j th) */
        cr += mr * c + mi * s;
        ci += mi * c - mr * s;
      }
      ir += cr;
      ii += ci;
      double vecx = ((double)p[0] + (double)o[0] / 2.0) - vx;
      double vecy = ((double)p[1] + (double)o[1] / 2.0) - vy;
      double vecz = ((double)p[2] + (double)o[2] / 2.0) - vz;
      m[0] += vecx * cr; m[1] += vecx * ci;
      m[2] += vecy * cr; m[3] += vecy * ci;
      m[4] += vecz * cr; m[5] += vecz * ci;
    }
    out[2 * t] = ir;
    out[2 * t + 1] = ii;
    if (mom)
      for (int q = 0; q < 6; ++q) mom[6 * t + q] = m[q];
  }
}

/*
 * imaging.cpp:60-139, full branch: FP32 complex sums, the phase formed in
 * double and narrowed to float for a complex<float> exponential, per TX and
 * per RX.  Here meas is rows x F with row = tx * R + rx (the layout of
 * or_bp_ref), ant holds one position per row.  Output divided by the TX
 * count as imaging.cpp:119 does.
 */
void or_bp_cpu_imaging(float *img, const float *xl, const float *yl,
                       const float *zl, const float *ant, const float *meas,
                       const float *rxo, float slope, float wavelength,
                       int nx, int ny, int nz, int n_tx, int R, int F,
                       int is_ti, const int64_t *vidx, int64_t nsel) {
  int64_t n = vidx ? nsel : (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < n; ++t) {
    int64_t lin = vidx ? vidx[t] : t;
    float vx, vy, vz;
    coords(xl, yl, zl, ny, nz, lin, &vx, &vy, &vz);
    float complex sum = 0.0f;
    for (int i = 0; i < n_tx; ++i) {
      const float *p = ant + 3 * (size_t)i * R;
      float ax = vx - p[0], ay = vy - p[1], az = vz - p[2];
      float fwd = sqrtf(ax * ax + ay * ay + az * az);
      if (is_ti) fwd = (float)((double)fwd + 0.15);
      for (int k = 0; k < R; ++k) {
        const float *o = rxo + 3 * k;
        float rxd = vx - (p[0] + o[0]), ryd = vy - (p[1] + o[1]), rzd = vz - (p[2] + o[2]);
        float back = sqrtf(rxd * rxd + ryd * ryd + rzd * rzd);
        float d = fwd + back;
        const float *row = meas + (size_t)2 * ((size_t)i * R + k) * F;
        float a0 = (float)(-2.0 * PI_D * (double)d / (double)wavelength);
        float complex e0 = cosf(a0) + I * sinf(a0);
        for (int j = 0; j < F; ++j) {
          float a1 = (float)(-2.0 * PI_D * (double)d * (double)j * (double)slope / C0);
          float complex e1 = cosf(a1) + I * sinf(a1);
          float complex m = row[2 * j] + I * row[2 * j + 1];
          sum += m * e1 * e0;
        }
      }
    }
    sum /= (float)n_tx;
    img[2 * t] = crealf(sum);
    img[2 * t + 1] = cimagf(sum);
  }
}

/*
 * sdf.py:180-222 for the given seeds (one per component, in visiting order).
 * rsdf/hops must hold V entries; they are initialised here (NaN / -1).
 * visited is shared across components as in sdf.py:142.
 */
static const int DXS[6] = {-1, 0, 1, 0, 0, 0};
static const int DYS[6] = {0, 1, 0, -1, 0, 0};
static const int DZS[6] = {0, 0, 0, 0, 1, -1};

int or_bfs_fifo(double *rsdf, int32_t *hops, const double *nrm,
                const int32_t *lab, const int64_t *seeds, int n_seeds,
                double dx, int nx, int ny, int nz) {
  int64_t V = (int64_t)nx * ny * nz;
  uint8_t *visited = calloc((size_t)V, 1);
  int64_t *q = malloc(sizeof(int64_t) * (size_t)(V > 0 ? V : 1));
  if (!visited || !q) {
    free(visited);
    free(q);
    return -1;
  }
  for (int64_t i = 0; i < V; ++i) {
    rsdf[i] = NAN;
    hops[i] = -1;
  }
  for (int c = 0; c < n_seeds; ++c) {
    int64_t s = seeds[c];
    int32_t label = lab[s];
    rsdf[s] = 0.0;
    if (hops[s] < 0) hops[s] = 0;
    int64_t head = 0, tail = 0;
    q[tail++] = s;
    visited[s] = 1;
    while (head < tail) {
      int64_t u = q[head++];
      int x = (int)(u / ((int64_t)ny * nz));
      int y = (int)((u / nz) % ny);
      int z = (int)(u % nz);
      double prev = rsdf[u];
      for (int d = 0; d < 6; ++d) {
        int ax = x + DXS[d], ay = y + DYS[d], az = z + DZS[d];
        if (ax < 0 || ay < 0 || az < 0 || ax >= nx || ay >= ny || az >= nz) continue;
        int64_t cidx = ((int64_t)ax * ny + ay) * nz + az;
        if (visited[cidx]) continue;
        if (isnan(nrm[3 * cidx])) continue;
        if (lab[cidx] != label) continue;
        q[tail++] = cidx;
        visited[cidx] = 1;
        hops[cidx] = hops[u] + 1;
        double t0 = ((double)DXS[d] * dx) * nrm[3 * cidx];
        double t1 = ((double)DYS[d] * dx) * nrm[3 * cidx + 1];
        double t2 = ((double)DZS[d] * dx) * nrm[3 * cidx + 2];
        rsdf[cidx] = prev + ((t0 + t1) + t2);
      }
    }
  }
  free(visited);
  free(q);
  return 0;
}

/* Li threshold on a = |img|^ex (img complex, interleaved), FP64 throughout:
 * x = a - min a; t = mean x; repeat { t_c = t; m_f = mean{x > t_c},
 * m_b = mean{x <= t_c}; stop if m_b == 0; t = (m_b - m_f)/(ln m_b - ln m_f) }
 * while t != t_c and fewer than max_iter updates of t were made.  Writes
 * t + min a and the number of updates; a constant image gives its value and
 * 0 updates.
 * Returns 0, or -1 on an allocation failure. */
int or_li_threshold(double *thr, int *iters, const float *img, int64_t n,
                    double ex, int max_iter) {
  double *a = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  if (!a) return -1;
  double lo = INFINITY, hi = -INFINITY, sum = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    a[i] = pow(hypot((double)img[2 * i], (double)img[2 * i + 1]), ex);
    if (a[i] < lo) lo = a[i];
    if (a[i] > hi) hi = a[i];
    sum += a[i];
  }
  *iters = 0;
  if (hi == lo) {
    *thr = lo;
    free(a);
    return 0;
  }
  double t_next = sum / (double)n - lo, t_curr = -1.0;
  int k = 0;
  while (fabs(t_next - t_curr) > 0.0 && k < max_iter) {
    t_curr = t_next;
    double sf = 0.0, nf = 0.0, sb = 0.0, nb = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double x = a[i] - lo;
      if (x > t_curr) {
        sf += x;
        nf += 1.0;
      } else {
        sb += x;
        nb += 1.0;
      }
    }
    const double mf = sf / nf, mb = sb / nb;
    if (mb == 0.0) break;
    t_next = (mb - mf) / (log(mb) - log(mf));
    ++k;
  }
  *thr = t_next + lo;
  *iters = k;
  free(a);
  return 0;
}

/* 6-connected labels: 0 background, 1..n by first appearance in a C-order
 * scan; each new component is flood-filled with a FIFO queue.  Returns n,
 * or -1 on an allocation failure. */
int or_ccl6(int32_t *labels, const uint8_t *mask, int nx, int ny, int nz) {
  const int64_t V = (int64_t)nx * ny * nz;
  int64_t *q = malloc(sizeof(int64_t) * (size_t)(V > 0 ? V : 1));
  if (!q) return -1;
  for (int64_t i = 0; i < V; ++i) labels[i] = 0;
  int n = 0;
  for (int64_t s = 0; s < V; ++s) {
    if (!mask[s] || labels[s]) continue;
    ++n;
    int64_t head = 0, tail = 0;
    q[tail++] = s;
    labels[s] = n;
    while (head < tail) {
      const int64_t u = q[head++];
      const int x = (int)(u / ((int64_t)ny * nz));
      const int y = (int)((u / nz) % ny);
      const int z = (int)(u % nz);
      for (int d = 0; d < 6; ++d) {
        const int ax = x + DXS[d], ay = y + DYS[d], az = z + DZS[d];
        if (ax < 0 || ay < 0 || az < 0 || ax >= nx || ay >= ny || az >= nz) continue;
        const int64_t c = ((int64_t)ax * ny + ay) * nz + az;
        if (!mask[c] || labels[c]) continue;
        labels[c] = n;
        q[tail++] = c;
      }
    }
  }
  free(q);
  return n;
}

/* ------------------------------------------------------------------ rays -- */
typedef struct {
  float fov_w, fov_h, bmin[3], bmax[3];
  int res[3];
  float eye[3], rxd[3];
  int width, height;
  float tau;
} RayGeom;

static float len3f(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

static void ray_dir(const RayGeom *g, int p, float *d) {
  const float pi = 3.14159265358979323846f;
  const float vfov = g->fov_h * (pi / 180.f), hfov = g->fov_w * (pi / 180.f);
  const float lw = len3f(g->rxd[0], g->rxd[1], g->rxd[2]);
  const float wx = g->rxd[0] / lw, wy = g->rxd[1] / lw, wz = g->rxd[2] / lw;
  const float cx = 0.0f * wz - 0.0f * wy, cy = 0.0f * wx - 1.0f * wz, cz = 1.0f * wy - 0.0f * wx;
  const float lc = len3f(cx, cy, cz);
  const float ux = cx / lc, uy = cy / lc, uz = cz / lc;
  const float vx = wy * uz - wz * uy, vy = wz * ux - wx * uz, vz = wx * uy - wy * ux;
  const int x = p % g->width, y = p / g->width;
  const float ha = (float)(-(double)hfov + (double)(hfov * 2) * ((double)(float)x + 0.5) /
                                               (double)(float)g->width);
  const float va = (float)(-(double)vfov + (double)(vfov * 2) * ((double)(float)y + 0.5) /
                                               (double)(float)g->height);
  const float px = tanf(ha), py = tanf(va);
  const float sx = px * ux + py * vx - wx, sy = px * uy + py * vy - wy, sz = px * uz + py * vz - wz;
  const float ls = len3f(sx, sy, sz);
  d[0] = sx / ls;
  d[1] = sy / ls;
  d[2] = sz / ls;
}

static float box_distance(const RayGeom *g, const float *v, float dx, float dy, float dz) {
  float dist = -1.f;
  dx = dx / len3f(dx, dy, dz);
  dy = dy / len3f(dx, dy, dz);
  dz = dz / len3f(dx, dy, dz);
  const float d[3] = {dx, dy, dz};
  for (int a = 0; a < 3; ++a) {
    if (dist == -1) {
      if ((v[a] < g->bmin[a]) && (d[a] > 0)) dist = (g->bmin[a] - v[a]) / d[a];
      if ((v[a] > g->bmax[a]) && (d[a] < 0)) dist = (g->bmax[a] - v[a]) / d[a];
    } else {
      if ((v[a] < g->bmin[a]) && (d[a] > 0)) dist = fmaxf(dist, (g->bmin[a] - v[a]) / d[a]);
      if ((v[a] > g->bmax[a]) && (d[a] < 0)) dist = fmaxf(dist, (g->bmax[a] - v[a]) / d[a]);
    }
  }
  return dist;
}

static float grid_at(const float *grid, const RayGeom *g, float x, float y, float z) {
  const float fz = (float)(int)z, fy = (float)(int)y, fx = (float)(int)x;
  const float f = fz + fy * (float)g->res[2] + fx * (float)g->res[2] * (float)g->res[1];
  int64_t i = (int64_t)(int)f;
  const int64_t V = (int64_t)g->res[0] * g->res[1] * g->res[2];
  if (i < 0) i = 0;
  if (i >= V) i = V - 1;
  return grid[i] - g->tau;
}

static float value_at(const float *grid, const RayGeom *g, const float *p, const float *d,
                      int first) {
  if (!(p[0] >= g->bmin[0] && p[0] <= g->bmax[0] && p[1] >= g->bmin[1] &&
        p[1] <= g->bmax[1] && p[2] >= g->bmin[2] && p[2] <= g->bmax[2])) {
    if (first) return box_distance(g, p, d[0], d[1], d[2]) + 0.00001f;
    return -1;
  }
  const float vs = (g->bmax[1] - g->bmin[1]) / (float)(g->res[1] - 1);
  float ix = floorf((p[0] - g->bmin[0]) / vs);
  float iy = floorf((p[1] - g->bmin[1]) / vs);
  float iz = floorf((p[2] - g->bmin[2]) / vs);
  if (ix == (g->bmax[0] - g->bmin[0]) / vs) ix = (g->bmax[0] - g->bmin[0]) / vs - 1;
  if (iy == (g->bmax[1] - g->bmin[1]) / vs) iy = (g->bmax[1] - g->bmin[1]) / vs - 1;
  if (iz == (g->bmax[2] - g->bmin[2]) / vs) iz = (g->bmax[2] - g->bmin[2]) / vs - 1;
  const float lim = -50.0f + 10;
  const float g000 = grid_at(grid, g, ix, iy, iz), g010 = grid_at(grid, g, ix, iy + 1, iz);
  const float g001 = grid_at(grid, g, ix, iy, iz + 1), g011 = grid_at(grid, g, ix, iy + 1, iz + 1);
  const float g100 = grid_at(grid, g, ix + 1, iy, iz), g110 = grid_at(grid, g, ix + 1, iy + 1, iz);
  const float g101 = grid_at(grid, g, ix + 1, iy, iz + 1);
  const float g111 = grid_at(grid, g, ix + 1, iy + 1, iz + 1);
  if (g000 < lim || g010 < lim || g001 < lim || g011 < lim || g100 < lim || g110 < lim ||
      g101 < lim || g111 < lim)
    return 0.005f;
  const float tx = (p[0] - (g->bmin[0] + ix * vs)) / vs;
  const float c01 = (1.f - tx) * g000 + tx * g100, c23 = (1.f - tx) * g010 + tx * g110;
  const float c45 = (1.f - tx) * g001 + tx * g101, c67 = (1.f - tx) * g011 + tx * g111;
  const float ty = (p[1] - (g->bmin[1] + iy * vs)) / vs;
  const float c0 = (1.f - ty) * c01 + ty * c23, c1 = (1.f - ty) * c45 + ty * c67;
  const float tz = (p[2] - (g->bmin[2] + iz * vs)) / vs;
  return (1.f - tz) * c0 + tz * c1;
}

static void intersect(const float *grid, const RayGeom *g, const float *d, float *hit,
                      float *vox) {
  const float vs = (g->bmax[1] - g->bmin[1]) / (float)(g->res[1] - 1);
  float depth = 0;
  int first = 1;
  for (int s = 0; s < 1000; ++s) {
    const float r[3] = {g->eye[0] + depth * d[0], g->eye[1] + depth * d[1],
                        g->eye[2] + depth * d[2]};
    const float dist = value_at(grid, g, r, d, first);
    first = 0;
    if (dist == -1) break;
    if (dist < vs && dist != 0.0f) {
      for (int a = 0; a < 3; ++a) {
        float v = floorf((r[a] - g->bmin[a]) / vs);
        if (v == (float)(g->res[a] - 1)) v = v - 1;
        vox[a] = v;
        hit[a] = r[a];
      }
      return;
    }
    depth += dist;
  }
  for (int a = 0; a < 3; ++a) hit[a] = vox[a] = -1.0f;
}

static void geom(RayGeom *g, int width, int height, float fov_w, float fov_h,
                 const float *bbox, const int *res, const float *eye, const float *rxd,
                 float tau) {
  g->fov_w = fov_w;
  g->fov_h = fov_h;
  for (int a = 0; a < 3; ++a) {
    g->bmin[a] = bbox[a];
    g->bmax[a] = bbox[3 + a];
    g->res[a] = res[a];
    g->eye[a] = eye[a];
    g->rxd[a] = rxd[a];
  }
  g->width = width;
  g->height = height;
  g->tau = tau;
}

void or_ray_matching(float *hit, float *vox, float *dir, const float *grid, int width,
                     int height, float fov_w, float fov_h, const float *bbox, const int *res,
                     const float *eye, const float *rxd, float tau) {
  RayGeom g;
  geom(&g, width, height, fov_w, fov_h, bbox, res, eye, rxd, tau);
  for (int p = 0; p < width * height; ++p) {
    ray_dir(&g, p, dir + 3 * (int64_t)p);
    intersect(grid, &g, dir + 3 * (int64_t)p, hit + 3 * (int64_t)p, vox + 3 * (int64_t)p);
  }
}

static int64_t torch_index(float v, int res) {
  int64_t i = (int64_t)v;
  if (i == -1) i = 0;
  if (i < 0) i += res;
  return i < 0 ? 0 : (i >= res ? res - 1 : i);
}

/* Per-ray FP64 round-trip distances (NaN: no hit) and the FP64 channel
 * sum_hits exp(+j 2 pi (f0 + j df) d / c) as interleaved doubles. */
void or_simulate_channel(double *chan, double *dist, const float *grid, float tau, int width,
                         int height, float fov_w, float fov_h, const float *bbox,
                         const int *res, const float *eye, const float *rxd,
                         const double *rx, float vsz, double f0, double df, int F,
                         int is_ti) {
  RayGeom g;
  geom(&g, width, height, fov_w, fov_h, bbox, res, eye, rxd, tau);
  const int64_t ny = res[1], nz = res[2];
  for (int j = 0; j < 2 * F; ++j) chan[j] = 0.0;
  for (int p = 0; p < width * height; ++p) {
    float d[3], hit[3], vox[3];
    ray_dir(&g, p, d);
    intersect(grid, &g, d, hit, vox);
    dist[p] = NAN;
    if (vox[0] == -1.0f) continue;
    int64_t ii[3];
    for (int a = 0; a < 3; ++a) ii[a] = torch_index(vox[a], res[a]);
#define AT(X, Y, Z)                                                                    \
  (grid[(((X) >= res[0] ? res[0] - 1 : (X)) * ny + ((Y) >= res[1] ? res[1] - 1 : (Y))) * nz + \
        ((Z) >= res[2] ? res[2] - 1 : (Z))] -                                          \
   tau)
    float t[3];
    for (int a = 0; a < 3; ++a) {
      const float vmin = g.bmin[a] + vox[a] * vsz;
      t[a] = (hit[a] - vmin) / vsz;
    }
    const int64_t x = ii[0], y = ii[1], z = ii[2];
    const float c01 = (1 - t[0]) * AT(x, y, z) + t[0] * AT(x + 1, y, z);
    const float c23 = (1 - t[0]) * AT(x, y + 1, z) + t[0] * AT(x + 1, y + 1, z);
    const float c45 = (1 - t[0]) * AT(x, y, z + 1) + t[0] * AT(x + 1, y, z + 1);
    const float c67 = (1 - t[0]) * AT(x, y + 1, z + 1) + t[0] * AT(x + 1, y + 1, z + 1);
#undef AT
    const float c0 = (1 - t[1]) * c01 + t[1] * c23;
    const float c1 = (1 - t[1]) * c45 + t[1] * c67;
    const float sv = (1 - t[2]) * c0 + t[2] * c1;
    double s = 0.0;
    for (int a = 0; a < 3; ++a) {
      const float q = hit[a] + d[a] * sv;
      const double e = rx[a] - (double)q;
      s += e * e;
    }
    double dd = sqrt(s) * 2;
    if (is_ti) dd += 0.15;
    dist[p] = dd;
    for (int j = 0; j < F; ++j) {
      const double ph = 2.0 * PI_D * dd * (f0 + j * df) / C0;
      chan[2 * j] += cos(ph);
      chan[2 * j + 1] += sin(ph);
    }
  }
}
