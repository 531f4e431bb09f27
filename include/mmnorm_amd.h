/*
 * mmnorm_amd.h -- C ABI of libmmnorm_hip.so, the MI355X (gfx950) image-formation
 * hot path of the mmNorm pipeline: SAR backprojection, the surface-normal vote,
 * and the RSDF breadth-first search.
 *
 * Conventions shared by every entry point:
 *   - All array arguments are DEVICE pointers owned by the caller.  Complex
 *     arrays are interleaved float pairs (re, im), i.e. the memory layout of
 *     a complex64 tensor.  The library allocates nothing; buffers it needs
 *     beyond inputs and outputs come in as a caller-provided workspace whose
 *     size the matching *_workspace_size() function returns.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Work is enqueued
 *     on it and the call returns without synchronising, except mmn_rsdf_bfs,
 *     which reads the frontier size back once per batch of levels.
 *   - Return value: 0 = success, < 0 = argument error (nothing launched),
 *     > 0 = a hipError_t from the runtime.  mmn_last_error() gives a
 *     thread-local message for the last non-zero status.
 *   - Voxel grids are separable: voxel (x, y, z) sits at
 *     (x_locs[x], y_locs[y], z_locs[z]) and has linear index
 *     (x * NY + y) * NZ + z (z fastest), as in imaging_gpu.cu:48-54,121.
 *   - Antenna rows: row i has TX position ant[i][0..2] and RX offset
 *     rx_off[i % R][0..2]; its measurement row is meas[i][0..F-1]
 *     (imaging_gpu.cu:65-76, the layout py_image_gpu.py:228-234 builds).
 *     The round-trip distance of row i to voxel v is
 *     |v - p_i| + |v - p_i - o_i| (+ 0.15 m when is_ti_radar != 0).
 */
#ifndef MMNORM_AMD_H
#define MMNORM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMN_ABI_VERSION 1

/* Backprojection variants accepted by mmn_bp_forward. */
#define MMN_BP_PLAIN 1 /* direct FP32 sincosf per (voxel, row, sample)          */
#define MMN_BP_REC   2 /* FP32 rotation (Horner) over samples, FP64 re-anchor
                          every K samples; plain FP32 sums within a row (its
                          K-sample blocks), Kahan-compensated FP32 sum over
                          rows                                                 */
#define MMN_BP_PLAIN_FAST 3 /* direct FP32: phase in revolutions, hardware sine
                               and cosine of its fraction                     */

/* Status codes below zero. */
#define MMN_ERR_ARG      (-1) /* bad size, count or enum value                */
#define MMN_ERR_NULL     (-2) /* a required pointer is NULL                    */
#define MMN_ERR_WS       (-3) /* workspace missing or too small                */
#define MMN_ERR_INPUT    (-4) /* input data violates a documented precondition */

int mmn_abi_version(void);
const char *mmn_last_error(void);

/*
 * mmn_cuda_image -- replaces the `cuda_image` kernel launch
 * (imaging_gpu.cu:11-123, launched at py_image_gpu.py:189-199) over the whole
 * grid.  Same arithmetic as that kernel: FP32 distance (sqrtf), FP64 phase
 * and FP64 cos/sin, FP32 accumulation in row order.  use_interpolated != 0
 * selects the range-bin branch (imaging_gpu.cu:90-105; meas rows hold
 * `samples` FFT bins), otherwise the full matched filter over `samples`
 * samples (imaging_gpu.cu:106-117).  Writes the raw sum, un-normalised,
 * to img[NX*NY*NZ] (complex).
 */
int mmn_cuda_image(float *img, const float *x_locs, const float *y_locs,
                   const float *z_locs, const float *ant, const float *meas,
                   const float *rx_off, float slope, float wavelength,
                   float fft_spacing, int nx, int ny, int nz, int num_rows,
                   int num_rx, int samples, int is_ti_radar,
                   int use_interpolated, void *stream);

/*
 * mmn_cuda_sn_est -- replaces the `cuda_sn_est` kernel launch
 * (surface_normal_est.cu:13-150, launched at py_sn_est_gpu.py:229-247).
 * For each of the n_valid voxels in vox_idx (linear indices; the reference
 * passes the same list as three index planes, py_sn_est_gpu.py:327-333)
 * writes the un-normalised vote  sum_i (p_i + o_i/2 - v) * Re(c_i conj(I_v))/|I_v|
 * to normals[n][3] (float), with c_i the interpolated range-bin contribution
 * of row i and I_v = img[vox] (complex).  Same arithmetic as the reference.
 */
int mmn_cuda_sn_est(float *normals, const int32_t *vox_idx, int64_t n_valid,
                    const float *img, const float *x_locs,
                    const float *y_locs, const float *z_locs, const float *ant,
                    const float *meas, const float *rx_off, float slope,
                    float wavelength, float fft_spacing, int nx, int ny,
                    int nz, int num_rows, int num_rx, int samples,
                    int is_ti_radar, void *stream);

/*
 * mmn_bp_forward -- full-frequency backprojection, the sum of
 * imaging_gpu.cu:106-117 written for MI355X:
 *     img[v] = sum_i sum_j meas[i][j] * exp(-j 2 pi (f0 + j*df) d_iv / c)
 * with c = 299792458 m/s.  The reference's phase -2 pi d/lambda - 2 pi d j
 * slope/c is this with f0 = c/lambda and df = slope.
 * If vox_idx != NULL the n_vox voxels it lists are computed and img/moment
 * are indexed by list position; otherwise all NX*NY*NZ voxels (n_vox
 * ignored).  If moment != NULL it also receives the normal-vote moment
 *     moment[v][k] = sum_i (p_i + o_i/2 - v)_k * c_iv      (complex, [n][3])
 * where c_iv is row i's contribution to img[v]; the un-normalised vote of
 * surface_normal_est.cu:128-144 is then Re(conj(I_v) moment[v]) / |I_v|
 * (mmn_sn_from_moment).  variant: MMN_BP_PLAIN, MMN_BP_PLAIN_FAST or
 * MMN_BP_REC; K (REC only)
 * is the FP64 re-anchor interval in samples, one of 8, 16, 32, 64, 128, 256
 * (or 0 = 64).  Up to 64 samples an interval is one FP32 Horner chain; 128 and
 * 256 are chains of 64 combined by a second Horner level in r^64, whose step
 * also comes from the FP64 range reduction (default lane shape only).
 * No reference counterpart for the fused moment or the variants; the
 * reference computes the vote only for the interpolated branch
 * (py_sn_est_gpu.py:404-405).
 * workspace (optional, may be NULL): when the voxel set gives fewer than
 * eight rounds of resident workgroups, lets every voxel block be split over
 * row segments (at least 32 rows each) whose partial sums are reduced in a
 * fixed order, so that the device stays full and each workgroup's window of
 * measurement rows stays in L2; mmn_bp_forward_workspace_size() gives
 * exactly the bytes of the split the launch chooses on the current device
 * (0 when it needs none; without a device it assumes MI355X: 256 CUs, 8
 * workgroups each).  A smaller workspace gives a coarser split or none.
 * Results are deterministic for a given workspace size and device.
 */
size_t mmn_bp_forward_workspace_size(int64_t n_vox, int num_rows, int with_moment);
int mmn_bp_forward(float *img, float *moment, const float *x_locs,
                   const float *y_locs, const float *z_locs,
                   const int32_t *vox_idx, int64_t n_vox, const float *ant,
                   const float *meas, const float *rx_off, double f0,
                   double df, int nx, int ny, int nz, int num_rows,
                   int num_rx, int samples, int is_ti_radar, int variant,
                   int K, void *workspace, size_t workspace_bytes,
                   void *stream);

/*
 * mmn_bp_profile -- the same full-frequency sum as mmn_bp_forward (same
 * arguments, image and moment), evaluated per (voxel, row) from a table of
 * the row's band-centred range profile
 *     Q_i(d) = sum_j meas[i][j] exp(-j 2 pi (j - (F-1)/2) df d / c)
 * sampled every h = min(c / (64 (F-1) |df|), c / (2 f_c)) metres and
 * interpolated by the cubic through the four nearest samples, times the
 * carrier exp(-j 2 pi f_c d / c), f_c = f0 + (F-1)/2 df, reduced in FP64 at
 * each table point.  The interpolation error is a few 1e-8 of
 * sum_j |meas[i][j]| per row.  bbox (host, 6 doubles: x, y, z minimum then
 * maximum) must contain every voxel; it sets the table length
 * M = ceil(2 |bbox diagonal| / h) + 6 per row.  workspace
 * is required: mmn_bp_profile_workspace_size() bytes (per-row tables, and
 * row-segment partials reduced in a fixed order).  No reference
 * counterpart: a faster evaluation of imaging_gpu.cu:106-117's sum.
 */
size_t mmn_bp_profile_workspace_size(int64_t n_vox, int num_rows, int samples, double f0,
                                     double df, const double *bbox, int with_moment);
int mmn_bp_profile(float *img, float *moment, const float *x_locs, const float *y_locs,
                   const float *z_locs, const int32_t *vox_idx, int64_t n_vox,
                   const float *ant, const float *meas, const float *rx_off, double f0,
                   double df, int nx, int ny, int nz, int num_rows, int num_rx,
                   int samples, int is_ti_radar, const double *bbox, void *workspace,
                   size_t workspace_bytes, void *stream);

/*
 * mmn_bp_profile_backward -- gradients of the full-frequency sum for a real
 * loss (the same quantities as mmn_bp_backward, arguments as there plus bbox
 * as in mmn_bp_profile), evaluated by groups of voxels instead of per sample:
 *     grad_ant[i]  = scale sum_v Im(conj(g_v) H_i(d_iv)) grad_p d_iv,
 *     H_i(d)       = sum_j k_j meas[i][j] exp(-j k_j d), k_j = 2 pi (f0 + j df)/c,
 *     grad_meas[i][j] = scale sum_v g_v exp(+j k_j d_iv).
 * Voxels are grouped (bricks of the grid, runs of a voxel list) so that
 * within a group d_iv = d0 + delta with |delta| small enough that the slow
 * factor exp(-+j 2 pi (j - (F-1)/2) df delta / c) is a short series in delta
 * (16 power-series terms for the position term, 12 Chebyshev terms for the
 * adjoint; truncation below 3e-8 of the row's sum of magnitudes).  One lane
 * per row; no atomics, deterministic.  Either output may be NULL.  bbox
 * (host, x, y, z minimum then maximum) sizes the groups; the call reads two
 * small values back to the host (the largest group radius, the RX offsets),
 * so it synchronises the stream.  workspace is required:
 * mmn_bp_profile_backward_workspace_size() bytes.  No reference
 * counterpart.
 */
size_t mmn_bp_profile_backward_workspace_size(int64_t n_vox, int num_rows, int samples,
                                              double f0, double df, const double *bbox);
int mmn_bp_profile_backward(float *grad_meas, float *grad_ant, const float *g, double scale,
                            const float *x_locs, const float *y_locs, const float *z_locs,
                            const int32_t *vox_idx, int64_t n_vox, const float *ant,
                            const float *meas, const float *rx_off, double f0, double df,
                            int nx, int ny, int nz, int num_rows, int num_rx, int samples,
                            int is_ti_radar, const double *bbox, void *workspace,
                            size_t workspace_bytes, void *stream);

/*
 * mmn_bp_interp -- the interpolated (range-bin) branch of imaging_gpu.cu:90-105
 * in MI355X form, plus the vote moment.  Per (voxel, row): the reference's
 * FP32 distance and bin floor(d / fft_spacing / 2) with the same operations
 * (so the same bins; rows with d > fft_spacing * samples are skipped), the
 * phasor exp(-j 2 pi d / wavelength) from the phase fraction formed with
 * 1/wavelength split into two floats (exact products by FMA) and the
 * hardware sine/cosine (the reference: FP64 cos/sin), and, if moment !=
 * NULL, moment[v][k] = sum_i (p_i + o_i/2 - v)_k * c_iv so that the vote of
 * surface_normal_est.cu:110-149 is mmn_sn_from_moment(img, moment).  meas
 * rows hold `samples` spectrum bins.  Voxel list and indexing as in
 * mmn_bp_forward.  No reference counterpart for the fused moment.
 */
int mmn_bp_interp(float *img, float *moment, const float *x_locs, const float *y_locs,
                  const float *z_locs, const int32_t *vox_idx, int64_t n_vox,
                  const float *ant, const float *meas, const float *rx_off,
                  float wavelength, float fft_spacing, int nx, int ny, int nz,
                  int num_rows, int num_rx, int samples, int is_ti_radar,
                  void *stream);

/*
 * mmn_sqrt_rn_check -- self-check of the square root mmn_bp_interp uses for
 * the reference's FP32 distances (the hardware square root and a neighbour
 * test, in place of sqrtf's longer sequence): over the float bit patterns
 * [first, first + count) (first + count <= 2^32) adds to result[0] the number
 * whose root differs from sqrtf's (NaNs compare equal) and lowers result[1]
 * to the smallest such pattern.  result is device memory the caller sets to
 * {0, ~0} first.  No reference counterpart: it pins that the distances, and
 * so the range bins, stay those of the reference's correctly rounded sqrtf.
 */
int mmn_sqrt_rn_check(uint64_t first, uint64_t count, unsigned long long *result, void *stream);

/*
 * mmn_sn_from_moment -- normals[n][3] = Re(conj(I) * moment[n][3]) / |I| with
 * I = img[vox_idx[n]] (img indexed by linear voxel index).  Completes the
 * full-frequency vote (the reference's weight, surface_normal_est.cu:138-139,
 * folded over rows by linearity).  A voxel with |I| = 0 gets NaN, as the
 * reference's division does.
 */
int mmn_sn_from_moment(float *normals, const float *img, const float *moment,
                       const int32_t *vox_idx, int64_t n_valid, void *stream);

/*
 * mmn_bp_backward -- adjoint of mmn_bp_forward (REC arithmetic) for a real
 * loss L of the image I = scale * img.  g = dL/dI in PyTorch's convention
 * (dL/dRe + i dL/dIm), complex [n] over the same voxels as the forward.
 * Writes
 *   grad_meas[i][j] = scale * sum_v g_v * exp(+j 2 pi f_j d_iv / c)      (complex [A][F])
 *   grad_ant[i][k]  = scale * sum_v Im(conj(g_v) Q_iv) * d(d_iv)/dp_ik    (float [A][3])
 * with Q_iv = sum_j (2 pi f_j / c) meas[i][j] exp(-j 2 pi f_j d_iv / c).
 * Either output may be NULL (not computed).  No reference counterpart:
 * the reference computes no gradients.  Deterministic: partial sums go to
 * the workspace and are reduced in a fixed order.  K (16 or 32, or 64 with
 * grad_meas NULL; 0 = 32) is the sample block of one wave and its re-anchor
 * interval: longer blocks cost less set-up and accumulate more rounding
 * (dL/dant within ~3e-6 of FP64 at K = 32, ~2e-5 at K = 64, on bistatic test
 * cases).
 */
size_t mmn_bp_backward_workspace_size(int num_rows, int samples, int64_t n_vox);
int mmn_bp_backward(float *grad_meas, float *grad_ant, const float *g,
                    double scale, const float *x_locs, const float *y_locs,
                    const float *z_locs, const int32_t *vox_idx,
                    int64_t n_vox, const float *ant, const float *meas,
                    const float *rx_off, double f0, double df, int nx,
                    int ny, int nz, int num_rows, int num_rx, int samples,
                    int is_ti_radar, int K, void *workspace,
                    size_t workspace_bytes, void *stream);

/*
 * mmn_rsdf_bfs -- replaces the per-component FIFO breadth-first search of
 * _construct_rsdf_components (sdf.py:180-222).  normals: float64 [V][3];
 * labels: int32 [V]; seeds: the n_seeds start voxels (linear indices), one
 * per component, in the order the reference visits components.  A voxel is
 * enterable from u when it is in bounds, not yet visited, has a non-NaN
 * normals[.][0] and labels equal to labels[u] (sdf.py:169-177).  Outputs
 * (all V entries written): rsdf float64 [V] (NaN where not reached, 0 at
 * seeds, otherwise rsdf[parent] + dot(dir * dx, normals[child]) with the
 * parent FIFO order gives, sdf.py:216-222), hops int32 [V] (BFS level, -1
 * where not reached).  Bit-exact with the reference's FIFO order.  Calls
 * hipStreamSynchronize on `stream` once per batch of levels.  Every seed
 * must lie in [0, V): the seeds are read on the device, which skips a seed
 * outside the grid and runs no level, and the call then returns
 * MMN_ERR_INPUT (rsdf and hops unspecified).
 */
size_t mmn_rsdf_bfs_workspace_size(int nx, int ny, int nz);
int mmn_rsdf_bfs(double *rsdf, int32_t *hops, const double *normals,
                 const int32_t *labels, const int32_t *seeds, int n_seeds,
                 double dx, int nx, int ny, int nz, void *workspace,
                 size_t workspace_bytes, void *stream);

/*
 * mmn_rsdf_bfs_noalloc -- mmn_rsdf_bfs with no workspace argument and no
 * allocation of any kind: its scratch is a device array of
 * MMN_RSDF_BFS_SCRATCH_BYTES (128 MiB) built into the library, one per GPU,
 * held from the first call on that GPU until the process ends.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (serving loops and harnesses that forbid hipMalloc).  Same
 * arguments, outputs, bit-exact results and stream synchronisation as
 * mmn_rsdf_bfs; the stream must belong to the current device.  A grid whose
 * scratch exceeds the array (mmn_rsdf_bfs_workspace_size(nx, ny, nz) >
 * MMN_RSDF_BFS_SCRATCH_BYTES, about 9.5 M voxels) returns MMN_ERR_ARG with
 * nothing launched.  All calls on one GPU share its array: calls that
 * overlap in time on one GPU (two host threads, or two streams) race and
 * must not be made.
 */
#define MMN_RSDF_BFS_SCRATCH_BYTES ((size_t)128 << 20)
int mmn_rsdf_bfs_noalloc(double *rsdf, int32_t *hops, const double *normals,
                         const int32_t *labels, const int32_t *seeds, int n_seeds,
                         double dx, int nx, int ny, int nz, void *stream);

/*
 * mmn_li_threshold -- replaces skimage.filters.threshold_li as called by
 * _separate_binary_image (sdf.py:46-70) on the image |img|^exponent.
 * img: complex [n] (finite values).  With a = |img|^exponent (FP64) and
 * x = a - min(a), iterates t <- (m_b - m_f) / (ln m_b - ln m_f) (m_f, m_b the
 * means of x above / not above t) from t = mean(x), stopping when t stops
 * changing, when m_b = 0, or after max_iter updates of t.  Writes the
 * threshold in the units of a (t + min a), the number of updates made (the
 * reference's iteration callbacks after the first) and whether the stopping
 * rule was met (host pointers).  A constant image
 * returns its value after 0 iterations.  If below != NULL it receives the
 * binary image of sdf.py:66, below[i] = a[i] < threshold (uint8 [n], device),
 * formed from the same a.  All sums are FP64 in a fixed order
 * (deterministic).  Synchronises `stream` once per 16 iterations.
 */
size_t mmn_li_threshold_workspace_size(int64_t n);
int mmn_li_threshold(double *threshold, int *iterations, int *converged,
                     uint8_t *below, const float *img, int64_t n,
                     double exponent, int max_iter, void *workspace,
                     size_t workspace_bytes, void *stream);

/*
 * mmn_ccl6 -- replaces cc3d.connected_components(mask, connectivity=6) of
 * _find_connected_components (sdf.py:89).  mask: uint8 [NX*NY*NZ] (non-zero
 * = foreground, z fastest).  labels: int32 [V], 0 on background and 1..n on
 * the n 6-connected components, numbered in the order a C-order scan first
 * meets them (scipy.ndimage.label's numbering; the component partition is
 * cc3d's).  If n_components != NULL the count is written there (host
 * pointer; the call then synchronises `stream`).
 */
size_t mmn_ccl6_workspace_size(int nx, int ny, int nz);
int mmn_ccl6(int32_t *labels, int32_t *n_components, const uint8_t *mask,
             int nx, int ny, int nz, void *workspace, size_t workspace_bytes,
             void *stream);

/*
 * mmn_label_sizes -- sizes[l] = number of the n_vox entries of labels equal
 * to l, for l = 1..n_labels, and sizes[0] = the number of all other entries
 * (int64 [n_labels + 1]).  The per-label areas of sdf.py:100 (ndimage.sum
 * of the mask over the labels).
 */
int mmn_label_sizes(int64_t *sizes, const int32_t *labels, int64_t n_vox,
                    int n_labels, void *stream);

/*
 * Ray marching and channel simulation on an RSDF isosurface (the consumer of
 * the RSDF).  Common arguments: `grid` float [NX*NY*NZ] (z fastest), the
 * SDF with undefined voxels at -50; bounding_box float[6] = (min x, y, z,
 * max x, y, z), the coordinates of the grid's corner voxels; grid_res
 * int[3]; eye float[3] the antenna position; rx_dir float[3] its boresight
 * (rays leave along -rx_dir through a width x height pinhole grid spanning
 * +-width_fov / +-height_fov degrees).  Host pointers for the small arrays.
 *
 * mmn_ray_matching -- replaces ray_tracing.ray_matching (ray_matching_cuda,
 * ray_tracing_kernel.cu:425-494): writes, per ray in pixel order
 * (x fastest), the rough hit, the index of its voxel's minimum corner (both
 * -1 for no hit) and the unit direction, each float [width*height][3].
 */
int mmn_ray_matching(float *intersection_pos, float *voxel_position,
                     float *directions, const float *grid, int width,
                     int height, float width_fov, float height_fov,
                     const float *bounding_box, const int *grid_res,
                     const float *eye, const float *rx_dir, void *stream);

/*
 * mmn_simulate_channel -- the channel of one antenna from the isosurface
 * grid - tau (_simulate_channels for one component,
 * isosurface_optimization.py:111-209): march every ray, refine the hit by
 * the interpolated value along the ray, d = 2 |rx_loc - hit| (+ 0.15 m if
 * is_ti_radar), and channel[j] = sum over hits of exp(+j 2 pi (f0 + j df) d / c)
 * (complex [samples]).  voxel_size is the refine step's voxel size (the
 * reference's z_locs[1] - z_locs[0]); rx_loc double[3] (host).  Workspace
 * from mmn_simulate_channel_workspace_size.
 */
size_t mmn_simulate_channel_workspace_size(int width, int height, int samples);
int mmn_simulate_channel(float *channel, const float *grid, float tau, int width,
                         int height, float width_fov, float height_fov,
                         const float *bounding_box, const int *grid_res,
                         const float *eye, const float *rx_dir,
                         const double *rx_loc, float voxel_size, double f0,
                         double df, int samples, int is_ti_radar,
                         void *workspace, size_t workspace_bytes, void *stream);

/*
 * mmn_simulate_channels -- n_sim channels of mmn_simulate_channel in one
 * launch: the search loop of isosurface_optimization.py:475-521 simulates
 * every (offset, antenna, component) of the grid search with one call each
 * (:111-209); here a batch shares one launch per kernel, each simulation
 * with its own antenna, direction, offset and component, and each channel
 * equals mmn_simulate_channel's for the same inputs bit for bit (the same
 * per-ray arithmetic, chunks and reduction order).  grid: n_comp components
 * of NX*NY*NZ floats each, one after another (device); comp[s] selects
 * simulation s's component and must lie in [0, n_comp).  Host arrays:
 * comp[n_sim], taus[n_sim], eyes[3 n_sim], rx_dirs[3 n_sim],
 * rx_locs[3 n_sim] (double), bounding_box[6], grid_res[3]; they are copied
 * before the call returns.  channels: complex [n_sim][samples] (device).
 * n_sim <= 65535.  Workspace from mmn_simulate_channels_workspace_size.
 */
size_t mmn_simulate_channels_workspace_size(int width, int height, int samples,
                                            int n_sim);
int mmn_simulate_channels(float *channels, const float *grid, int n_comp,
                          const int32_t *comp, const float *taus, const float *eyes,
                          const float *rx_dirs, const double *rx_locs, int n_sim,
                          int width, int height, float width_fov, float height_fov,
                          const float *bounding_box, const int *grid_res,
                          float voxel_size, double f0, double df, int samples,
                          int is_ti_radar, void *workspace, size_t workspace_bytes,
                          void *stream);

/*
 * mmn_ray_prefix / mmn_simulate_channels_prefixed -- the isosurface offset
 * search (isosurface_optimization.py:475-521) simulates the same antennas
 * and components for every offset tau.  A ray's march does not depend on
 * tau while the ray is outside the box or in a cell with an undefined corner
 * (grid - tau < UNDEF_VAL + 10 only turns true as tau grows, so a cell
 * undefined at the search's smallest offset tau_lo is undefined at every
 * tau >= tau_lo, and its step is MIN_STEP).  mmn_ray_prefix marches each ray
 * of n_sim simulations (comp, eyes, rx_dirs as for mmn_simulate_channels) up
 * to its first step in a cell fully defined at tau_lo, records that step, and
 * lists the rays that reach one or hit before it ("live" rays); it writes
 * each simulation's live-ray count to live_counts[n_sim] (host) and
 * synchronises the stream.  mmn_simulate_channels_prefixed then gives the
 * channels of mmn_simulate_channels for offsets taus[s] >= tau_lo, bit for
 * bit, marching and summing the live rays only.  The prefix buffer
 * (mmn_ray_prefix_size bytes, device) belongs to one set of simulations:
 * pass the same comp, eyes, rx_dirs, grid and geometry to both calls.
 */
size_t mmn_ray_prefix_size(int width, int height, int n_sim);
int mmn_ray_prefix(void *prefix, size_t prefix_bytes, int *live_counts, const float *grid,
                   int n_comp, const int32_t *comp, const float *eyes, const float *rx_dirs,
                   int n_sim, int width, int height, float width_fov, float height_fov,
                   const float *bounding_box, const int *grid_res, float tau_lo,
                   void *stream);
size_t mmn_simulate_channels_prefixed_workspace_size(int width, int height, int samples,
                                                     int n_sim, int live_max);
int mmn_simulate_channels_prefixed(float *channels, const void *prefix, size_t prefix_bytes,
                                   const int *live_counts, const float *grid, int n_comp,
                                   const int32_t *comp, const float *taus, const float *eyes,
                                   const float *rx_dirs, const double *rx_locs, int n_sim,
                                   int width, int height, float width_fov,
                                   float height_fov, const float *bounding_box,
                                   const int *grid_res, float voxel_size, double f0,
                                   double df, int samples, int is_ti_radar, float tau_lo,
                                   void *workspace, size_t workspace_bytes, void *stream);

/*
 * mmn_channel_loss -- loss_fn of isosurface_optimization.py:211-246 for
 * n_pairs (simulated, measured) channel pairs (complex [n_pairs][samples]
 * each): the DFT of each zero-padded to n_fft, bins [bin_lo, bin_hi) (the
 * measured profile zero from bin_hi_target on), each profile scaled to unit
 * L2 norm, loss[p] = sum over the bins of | |sim| - |measured| |.  At most
 * 1024 bins.
 */
int mmn_channel_loss(float *loss, const float *sim, const float *target,
                     int n_pairs, int samples, int n_fft, int bin_lo, int bin_hi,
                     int bin_hi_target, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MMNORM_AMD_H */
