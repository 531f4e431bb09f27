"""CPU oracle for the image-formation hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker or the timed CPU baseline.  The product
(package ``mmnorm_amd``) never calls it.

Two independent restatements of the reference's arithmetic live here:

* ``liboracle.so`` (oracle/bp_oracle.c, built by oracle/Makefile), fast,
  OpenMP over voxels -- or_bp_ref (imaging_gpu.cu:48-122), or_sn_ref
  (surface_normal_est.cu:64-149), or_bp_f64 (FP64 oracle O64 of the full
  sum + vote moment), or_bp_cpu_imaging (imaging.cpp:60-139) and
  or_bfs_fifo (sdf.py:180-222);
* pure-Python loops below (``py_*``) for tiny inputs, written line by line
  from the same reference lines, used to cross-check the C code.

Pins (tests/test_oracle.py, tests/test_oracle_interp.py): closed forms
(full-branch point-scatterer focus I = rho*A*F; interpolated branch: bins
and the skip rule on hand-constructed distances, the geometric-sum focus of
an FFT'd point scatterer; vote: a single row's direction and weight, a
symmetric four-RX aperture's normal (0,0,1); constant-normal RSDF =
dx*(k-k0), BFS hops = Manhattan distance in a full box), invariants
(linearity, translation, row permutation, label renumbering) and brute
force (the py_* loops).  The
reference ships no golden vectors, tests or fixtures for this path.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from collections import deque

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
C0 = 2.99792458e8

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        p, i32, i64, f32, f64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                 ctypes.c_float, ctypes.c_double)
        L.or_bp_ref.argtypes = [p, p, p, p, p, p, p, f32, f32, f32, i32, i32, i32,
                                i32, i32, i32, i32, i32, p, i64]
        L.or_sn_ref.argtypes = [p, p, i64, p, p, p, p, p, p, p, f32, f32, i32, i32,
                                i32, i32, i32, i32]
        L.or_bp_f64.argtypes = [p, p, p, p, p, p, p, p, f64, f64, i32, i32, i32,
                                i32, i32, i32, i32, p, i64]
        L.or_bp_cpu_imaging.argtypes = [p, p, p, p, p, p, p, f32, f32, i32, i32, i32,
                                        i32, i32, i32, i32, p, i64]
        L.or_bfs_fifo.argtypes = [p, p, p, p, p, i32, f64, i32, i32, i32]
        L.or_bfs_fifo.restype = i32
        L.or_li_threshold.argtypes = [p, p, p, i64, f64, i32]
        L.or_li_threshold.restype = i32
        L.or_ccl6.argtypes = [p, p, i32, i32, i32]
        L.or_ccl6.restype = i32
        L.or_ray_matching.argtypes = [p, p, p, p, i32, i32, f32, f32, p, p, p, p, f32]
        L.or_ray_matching.restype = None
        L.or_simulate_channel.argtypes = [p, p, p, f32, i32, i32, f32, f32, p, p, p, p, p,
                                          f32, f64, f64, i32, i32]
        L.or_simulate_channel.restype = None
        for fn in (L.or_bp_ref, L.or_sn_ref, L.or_bp_f64, L.or_bp_cpu_imaging):
            fn.restype = None
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _ptr(a):
    return None if a is None else a.ctypes.data


def _sel(vidx):
    if vidx is None:
        return None, 0
    v = np.ascontiguousarray(vidx, dtype=np.int64)
    return v, len(v)


def bp_ref(xl, yl, zl, ant, meas, rxo, slope, wavelength, fft_spacing=1.0,
           is_ti=False, interp=False, vidx=None):
    """Raw sum of `cuda_image` (imaging_gpu.cu:48-122), complex64.

    ant: (rows, 3) positions per row; meas: (rows, S) complex; rxo: (R, 3).
    """
    xl, yl, zl, ant, rxo = map(_f32, (xl, yl, zl, ant, rxo))
    meas = _c64(meas)
    rows, S = meas.shape
    v, n = _sel(vidx)
    nv = n if v is not None else len(xl) * len(yl) * len(zl)
    out = np.zeros(nv, np.complex64)
    lib().or_bp_ref(_ptr(out), _ptr(xl), _ptr(yl), _ptr(zl), _ptr(ant), _ptr(meas),
                    _ptr(rxo), float(np.float32(slope)), float(np.float32(wavelength)),
                    float(np.float32(fft_spacing)), len(xl), len(yl), len(zl), rows,
                    len(rxo), S, int(bool(is_ti)), int(bool(interp)), _ptr(v), n)
    return out


def sn_ref(vidx, img, xl, yl, zl, ant, meas, rxo, wavelength, fft_spacing, is_ti=False):
    """Raw vote of `cuda_sn_est` (surface_normal_est.cu:64-149), float32 (n, 3)."""
    xl, yl, zl, ant, rxo = map(_f32, (xl, yl, zl, ant, rxo))
    meas, img = _c64(meas), _c64(np.ravel(img))
    rows, S = meas.shape
    v = np.ascontiguousarray(vidx, dtype=np.int64)
    out = np.zeros((len(v), 3), np.float32)
    lib().or_sn_ref(_ptr(out), _ptr(v), len(v), _ptr(img), _ptr(xl), _ptr(yl), _ptr(zl),
                    _ptr(ant), _ptr(meas), _ptr(rxo), float(np.float32(wavelength)),
                    float(np.float32(fft_spacing)), len(yl), len(zl), rows, len(rxo), S,
                    int(bool(is_ti)))
    return out


def bp_f64(xl, yl, zl, ant, meas, rxo, f0, df, is_ti=False, vidx=None, moment=False):
    """FP64 oracle O64 of the full sum (and vote moment), from FP32-rounded inputs.

    Returns complex128 (n,) [, complex128 (n, 3)].
    """
    xl, yl, zl, ant, rxo = map(_f32, (xl, yl, zl, ant, rxo))
    meas = _c64(meas)
    rows, F = meas.shape
    v, n = _sel(vidx)
    nv = n if v is not None else len(xl) * len(yl) * len(zl)
    out = np.zeros(nv, np.complex128)
    mom = np.zeros((nv, 3), np.complex128) if moment else None
    lib().or_bp_f64(_ptr(out), _ptr(mom), _ptr(xl), _ptr(yl), _ptr(zl), _ptr(ant),
                    _ptr(meas), _ptr(rxo), float(f0), float(df), len(xl), len(yl),
                    len(zl), rows, len(rxo), F, int(bool(is_ti)), _ptr(v), n)
    return (out, mom) if moment else out


def np_bp_full(xl, yl, zl, ant, meas, rxo, slope, wavelength, vidx, dtype=np.float64,
               chunk=4):
    """NumPy restatement of the full branch of imaging_gpu.cu:48-122 with the
    vote moment sum_i (p_i + o_i/2 - v) c_i, vectorised over chunks of the
    listed voxels (one core: NumPy's ufuncs are single-threaded).  dtype
    float64: distance, phase -2 pi d / lambda - 2 pi d j slope / c and exp in
    FP64; float32: all in FP32 (complex64 exp, FP32 sums), the precision of
    the reference's CPU imaging (imaging.cpp:108-112).  The CPU-1 baseline
    bench.py times.  Returns (img (n,), moment (n, 3)) of the dtype's complex."""
    ft = np.dtype(dtype).type
    ct = np.complex128 if ft is np.float64 else np.complex64
    xl, yl, zl = (np.asarray(a, np.float32).astype(ft) for a in (xl, yl, zl))
    ant = np.asarray(ant, np.float32).astype(ft)
    rxo = np.asarray(rxo, np.float32).astype(ft).reshape(-1, 3)
    meas = np.asarray(meas, np.complex64).astype(ct)
    A, F = meas.shape
    o = rxo[np.arange(A) % len(rxo)]
    vidx = np.asarray(vidx, np.int64)
    ny, nz = len(yl), len(zl)
    ix, r = np.divmod(vidx, ny * nz)
    iy, iz = np.divmod(r, nz)
    vox = np.stack([xl[ix], yl[iy], zl[iz]], 1)
    two_pi = ft(2 * math.pi)
    kj = (two_pi / ft(wavelength) + two_pi * np.arange(F, dtype=ft) * ft(slope) / ft(C0))
    va = ant + o / ft(2)
    img = np.empty(len(vidx), ct)
    mom = np.empty((len(vidx), 3), ct)
    for lo in range(0, len(vidx), chunk):
        v = vox[lo:lo + chunk]                                   # (n, 3)
        a = v[:, None, :] - ant[None, :, :]                      # (n, A, 3)
        d = np.sqrt(np.sum(a * a, -1)) + np.sqrt(np.sum((a - o[None]) ** 2, -1))
        ph = np.exp(-1j * (d[:, :, None] * kj[None, None, :]).astype(ft)).astype(ct)
        c = np.sum(ph * meas[None], -1)                          # (n, A) row contributions
        img[lo:lo + chunk] = c.sum(-1)
        mom[lo:lo + chunk] = np.einsum("nak,na->nk", (va[None] - v[:, None, :]).astype(ct), c)
    return img, mom


def bp_cpu_imaging(xl, yl, zl, ant_tx, meas, rxo, slope, wavelength, is_ti=False,
                   vidx=None):
    """imaging.cpp:60-139 full branch (CPU baseline formulation), complex64.

    ant_tx: (n_tx, 3); meas: (n_tx * R, F) with row = tx * R + rx.
    """
    xl, yl, zl, rxo = map(_f32, (xl, yl, zl, rxo))
    R = len(rxo)
    ant_rows = _f32(np.repeat(np.asarray(ant_tx, np.float32), R, axis=0))
    meas = _c64(meas)
    F = meas.shape[1]
    v, n = _sel(vidx)
    nv = n if v is not None else len(xl) * len(yl) * len(zl)
    out = np.zeros(nv, np.complex64)
    lib().or_bp_cpu_imaging(_ptr(out), _ptr(xl), _ptr(yl), _ptr(zl), _ptr(ant_rows),
                            _ptr(meas), _ptr(rxo), float(np.float32(slope)),
                            float(np.float32(wavelength)), len(xl), len(yl), len(zl),
                            len(ant_tx), R, F, int(bool(is_ti)), _ptr(v), n)
    return out


def bfs_fifo(normals, labels, seeds, dx):
    """sdf.py:180-222 from given seeds; returns (rsdf float64 V, hops int32 V)."""
    nx, ny, nz = labels.shape
    nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
    lab = np.ascontiguousarray(labels, dtype=np.int32).ravel()
    s = np.ascontiguousarray(seeds, dtype=np.int64)
    V = nx * ny * nz
    rsdf = np.empty(V, np.float64)
    hops = np.empty(V, np.int32)
    rc = lib().or_bfs_fifo(_ptr(rsdf), _ptr(hops), _ptr(nrm), _ptr(lab), _ptr(s), len(s),
                           float(dx), nx, ny, nz)
    if rc != 0:
        raise MemoryError("or_bfs_fifo: allocation failed")
    return rsdf.reshape(nx, ny, nz), hops.reshape(nx, ny, nz)


def li_threshold(img, exponent, max_iter=1000):
    """Li threshold of |img|**exponent (FP64, sequential sums): (threshold, iterations)."""
    z = np.ascontiguousarray(np.asarray(img, np.complex64).ravel())
    thr = ctypes.c_double()
    it = ctypes.c_int()
    rc = lib().or_li_threshold(ctypes.addressof(thr), ctypes.addressof(it), _ptr(z), z.size,
                               float(exponent), int(max_iter))
    if rc != 0:
        raise MemoryError("or_li_threshold: allocation failed")
    return thr.value, it.value


def ccl6(mask):
    """6-connected labels (0 background, 1..n by first appearance in C order), n."""
    m = np.ascontiguousarray(np.asarray(mask, dtype=bool).astype(np.uint8))
    nx, ny, nz = m.shape
    lab = np.empty(m.shape, np.int32)
    n = lib().or_ccl6(_ptr(lab), _ptr(m), nx, ny, nz)
    if n < 0:
        raise MemoryError("or_ccl6: allocation failed")
    return lab, n


def _ray_args(grid, bbox, eye, rx_dir):
    g = np.ascontiguousarray(grid, dtype=np.float32)
    bb = np.ascontiguousarray(bbox, dtype=np.float32)
    res = np.ascontiguousarray(g.shape, dtype=np.int32)
    e = np.ascontiguousarray(eye, dtype=np.float32)
    r = np.ascontiguousarray(rx_dir, dtype=np.float32)
    return g, bb, res, e, r


def ray_matching(grid, width, height, width_fov, height_fov, bbox, eye, rx_dir, tau=0.0):
    """ray_tracing_kernel.cu:60-421: (hit, voxel, direction), each (width*height, 3)."""
    g, bb, res, e, r = _ray_args(grid, bbox, eye, rx_dir)
    n = width * height
    hit, vox, d = (np.empty((n, 3), np.float32) for _ in range(3))
    lib().or_ray_matching(_ptr(hit), _ptr(vox), _ptr(d), _ptr(g), width, height,
                          float(width_fov), float(height_fov), _ptr(bb), _ptr(res), _ptr(e),
                          _ptr(r), float(tau))
    return hit, vox, d


def simulate_channel(grid, tau, width, height, width_fov, height_fov, bbox, eye, rx_dir,
                     rx_loc, voxel_size, f0, df, F, is_ti=False):
    """isosurface_optimization.py:111-209 for one component, FP64 exp per hit and
    sample: (channel complex128 (F,), distances float64 (width*height,), NaN = no hit)."""
    g, bb, res, e, r = _ray_args(grid, bbox, eye, rx_dir)
    rx = np.ascontiguousarray(rx_loc, dtype=np.float64)
    ch = np.empty(2 * F, np.float64)
    dist = np.empty(width * height, np.float64)
    lib().or_simulate_channel(_ptr(ch), _ptr(dist), _ptr(g), float(tau), width, height,
                              float(width_fov), float(height_fov), _ptr(bb), _ptr(res), _ptr(e),
                              _ptr(r), _ptr(rx), float(np.float32(voxel_size)), float(f0),
                              float(df), int(F), int(bool(is_ti)))
    return ch[0::2] + 1j * ch[1::2], dist


def py_ray_dir(width, height, width_fov, height_fov, rx_dir, p):
    """GenerateRay (ray_tracing_kernel.cu:60-127) for ray p with NumPy float32 scalars."""
    f = np.float32
    pi = f(3.14159265358979323846)
    vfov, hfov = f(height_fov) * (pi / f(180)), f(width_fov) * (pi / f(180))
    rd = [f(v) for v in rx_dir]

    def ln(x, y, z):
        return np.sqrt(x * x + y * y + z * z, dtype=np.float32)
    lw = ln(*rd)
    w = [rd[0] / lw, rd[1] / lw, rd[2] / lw]
    c = [f(0) * w[2] - f(0) * w[1], f(0) * w[0] - f(1) * w[2], f(1) * w[1] - f(0) * w[0]]
    lc = ln(*c)
    u = [c[0] / lc, c[1] / lc, c[2] / lc]
    v = [w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]]
    x, y = p % width, p // width
    ha = f(-float(hfov) + float(hfov * f(2)) * (float(f(x)) + 0.5) / float(f(width)))
    va = f(-float(vfov) + float(vfov * f(2)) * (float(f(y)) + 0.5) / float(f(height)))
    px, py = np.tan(ha, dtype=np.float32), np.tan(va, dtype=np.float32)
    s = [px * u[a] + py * v[a] - w[a] for a in range(3)]
    ls = ln(*s)
    return np.array([s[0] / ls, s[1] / ls, s[2] / ls], np.float32)


def py_li_threshold(img, exponent, max_iter=1000):
    """The same iteration with NumPy float64 sums (pairwise): tiny inputs."""
    a = np.abs(np.asarray(img, np.complex64).astype(np.complex128).ravel()) ** exponent
    lo = a.min()
    if a.max() == lo:
        return float(lo), 0
    x = a - lo
    t_next, t_curr, k = x.mean(), -1.0, 0
    while abs(t_next - t_curr) > 0 and k < max_iter:
        t_curr = t_next
        fg = x > t_curr
        mf, mb = x[fg].mean(), x[~fg].mean()
        if mb == 0:
            break
        t_next = (mb - mf) / (np.log(mb) - np.log(mf))
        k += 1
    return float(t_next + lo), k


class LiAbort(Exception):
    """The reference's callback limit was exceeded (sdf.py:53-58)."""


def py_skimage_li(image, max_callbacks=None):
    """skimage.filters.threshold_li of scikit-image 0.24.0 -- the version the
    reference pins (requirements.txt:25) and calls at sdf.py:60-70, not
    installed here -- restated from its published algorithm for a float
    image, in the image's own dtype with NumPy's own mean and log (a float32
    image keeps float32 means, NumPy's pairwise sums, as in the reference's
    run):

      NaNs dropped; a constant image returns its value; infinities dropped;
      x = image - min; tolerance = half the smallest gap between distinct
      values of x; t_next = mean(x), t_curr = -2 tolerance; the callback is
      called once on the initial guess; while |t_next - t_curr| > tolerance:
      t_curr = t_next, m_f = mean(x > t_curr), m_b = mean(x <= t_curr),
      stop if m_b == 0, t_next = (m_b - m_f) / (log m_b - log m_f), callback;
      threshold = t_next + min.

    Returns (threshold, callbacks).  With ``max_callbacks`` (the reference
    passes a callback that raises on call 1001) raises LiAbort on the call
    past it."""
    image = np.asarray(image)
    image = image[~np.isnan(image)]
    if image.size == 0:
        return np.nan, 0
    if np.all(image == image.flat[0]):
        return image.flat[0], 0
    image = image[np.isfinite(image)]
    if image.size == 0:
        return 0.0, 0
    image_min = np.min(image)
    image = image - image_min
    tolerance = np.min(np.diff(np.unique(image))) / 2
    t_next = np.mean(image)
    t_curr = -2 * tolerance
    calls = 1
    while abs(t_next - t_curr) > tolerance:
        t_curr = t_next
        foreground = image > t_curr
        mean_fore = np.mean(image[foreground])
        mean_back = np.mean(image[~foreground])
        if mean_back == 0:
            break
        t_next = (mean_back - mean_fore) / (np.log(mean_back) - np.log(mean_fore))
        calls += 1
        if max_callbacks is not None and calls > max_callbacks:
            raise LiAbort(calls)
    return t_next + image_min, calls


def skimage_li_mask(img, exponent):
    """The reference's binary image for one exponent (sdf.py:62-67):
    |img|**exponent in the image's float32 and the restated threshold_li;
    returns (mask |img|**exp < threshold, threshold, callbacks, |img|**exp)."""
    abs_img = np.abs(np.asarray(img, np.complex64)) ** exponent
    thr, calls = py_skimage_li(abs_img)
    return abs_img < thr, thr, calls, abs_img


# ----------------------------------------------------------------------------
# Seeds as _construct_rsdf_components picks them (sdf.py:180-191).
# ----------------------------------------------------------------------------
def rsdf_seeds(all_normals, coords_full, labeled, valid_labels):
    """Linear index of each component's start voxel, in valid_labels order."""
    seeds = []
    ok = ~np.isnan(all_normals[:, :, :, 0])
    for comp in valid_labels:
        mask = np.logical_and(labeled == comp, ok)
        idxs = np.argwhere(mask)
        sel = coords_full[mask, :]
        mean = np.mean(sel, axis=0)[None, :]
        k = int(np.argmin(np.linalg.norm(sel - mean, axis=1)))
        i0, j0, k0 = idxs[k]
        seeds.append(np.ravel_multi_index((i0, j0, k0), labeled.shape))
    return np.asarray(seeds, np.int64)


# ----------------------------------------------------------------------------
# Pure-Python restatements for tiny inputs (cross-checks of the C code).
# ----------------------------------------------------------------------------
def py_bp_ref(xl, yl, zl, ant, meas, rxo, slope, wavelength, fft_spacing=1.0,
              is_ti=False, interp=False):
    """imaging_gpu.cu:48-122 with numpy float32/float64 scalars, tiny inputs only."""
    f32, f64 = np.float32, np.float64
    xl, yl, zl, ant, rxo = map(_f32, (xl, yl, zl, ant, rxo))
    meas = _c64(meas)
    rows, S = meas.shape
    R = len(rxo)
    slope, wavelength, fs = f32(slope), f32(wavelength), f32(fft_spacing)
    out = np.zeros(len(xl) * len(yl) * len(zl), np.complex64)
    pi = f64(math.pi)
    for x in range(len(xl)):
        for y in range(len(yl)):
            for z in range(len(zl)):
                sr, si = f32(0), f32(0)
                for i in range(rows):
                    dx, dy, dz = xl[x] - ant[i, 0], yl[y] - ant[i, 1], zl[z] - ant[i, 2]
                    o = rxo[i % R]
                    fwd = np.sqrt(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz), dtype=f32)
                    bx, by, bz = f32(dx - o[0]), f32(dy - o[1]), f32(dz - o[2])
                    back = np.sqrt(f32(f32(bx * bx) + f32(by * by)) + f32(bz * bz), dtype=f32)
                    d = f32(fwd + back)
                    if is_ti:
                        d = f32(f64(d) + 0.15)
                    if interp:
                        if d < 0 or d > f32(fs * f32(S)):
                            continue
                        b = int(np.floor(f32(f32(d / fs) / f32(2))))
                        m = meas[i, b]
                        th = f64(-2.0) * pi * f64(d) / f64(wavelength)
                        pr, pim = f32(np.cos(th)), f32(np.sin(th))
                        sr = f32(sr + f32(f32(m.real * pr) - f32(m.imag * pim)))
                        si = f32(si + f32(f32(m.real * pim) + f32(m.imag * pr)))
                    else:
                        for j in range(S):
                            m = meas[i, j]
                            th = (f64(-2.0) * pi * f64(d) / f64(wavelength)
                                  + f64(-2.0) * pi * f64(d) * f64(j) * f64(slope) / C0)
                            pr, pim = f32(np.cos(th)), f32(np.sin(th))
                            sr = f32(sr + f32(f32(m.real * pr) - f32(m.imag * pim)))
                            si = f32(si + f32(f32(m.real * pim) + f32(m.imag * pr)))
                out[(x * len(yl) + y) * len(zl) + z] = complex(sr, si)
    return out


def py_bfs_fifo(all_normals, labeled, seeds, delta_x):
    """sdf.py:195-222 literally (deque, isValid, np.dot), from given seeds."""
    rsdf = np.full(labeled.shape, np.nan)
    hops = np.full(labeled.shape, -1, np.int32)
    visited = np.zeros_like(labeled)
    dX = [-1, 0, 1, 0, 0, 0]
    dY = [0, 1, 0, -1, 0, 0]
    dZ = [0, 0, 0, 0, 1, -1]

    def valid(x, y, z, lab):
        if x < 0 or y < 0 or z < 0 or x >= visited.shape[0] or y >= visited.shape[1] \
                or z >= visited.shape[2]:
            return False
        if visited[x, y, z]:
            return False
        if np.isnan(all_normals[x, y, z, 0]):
            return False
        return labeled[x, y, z] == lab

    for s in seeds:
        i0, j0, k0 = np.unravel_index(int(s), labeled.shape)
        lab = labeled[i0, j0, k0]
        rsdf[i0, j0, k0] = 0
        if hops[i0, j0, k0] < 0:
            hops[i0, j0, k0] = 0
        q = deque([(i0, j0, k0)])
        visited[i0, j0, k0] = True
        while q:
            i, j, k = q.popleft()
            prev = rsdf[i, j, k]
            for d in range(6):
                a, b, c = i + dX[d], j + dY[d], k + dZ[d]
                vec = np.array([dX[d], dY[d], dZ[d]])
                if valid(a, b, c, lab):
                    q.append((a, b, c))
                    visited[a, b, c] = True
                    hops[a, b, c] = hops[i, j, k] + 1
                    rsdf[a, b, c] = prev + np.dot(vec * delta_x, all_normals[a, b, c])
    return rsdf, hops


def py_sn_ref(vidx, img, xl, yl, zl, ant, meas, rxo, wavelength, fft_spacing, is_ti=False):
    """surface_normal_est.cu:64-149 with numpy float32/float64 scalars, tiny inputs only."""
    f32, f64 = np.float32, np.float64
    xl, yl, zl, ant, rxo = map(_f32, (xl, yl, zl, ant, rxo))
    meas, img = _c64(meas), _c64(np.ravel(img))
    rows, S = meas.shape
    R = len(rxo)
    wavelength, fs = f32(wavelength), f32(fft_spacing)
    ny, nz = len(yl), len(zl)
    out = np.zeros((len(vidx), 3), np.float32)
    pi = f64(math.pi)
    for t, lin in enumerate(vidx):
        x, r = divmod(int(lin), ny * nz)
        y, z = divmod(r, nz)
        vx, vy, vz = xl[x], yl[y], zl[z]
        ir, ii = f32(img[lin].real), f32(img[lin].imag)
        mag = np.sqrt(f32(f32(ir * ir) + f32(ii * ii)), dtype=f32)
        sn = [f32(0), f32(0), f32(0)]
        for i in range(rows):
            p, o = ant[i], rxo[i % R]
            dx, dy, dz = f32(vx - p[0]), f32(vy - p[1]), f32(vz - p[2])
            fwd = np.sqrt(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz), dtype=f32)
            bx, by, bz = f32(dx - o[0]), f32(dy - o[1]), f32(dz - o[2])
            back = np.sqrt(f32(f32(bx * bx) + f32(by * by)) + f32(bz * bz), dtype=f32)
            d = f32(fwd + back)
            if is_ti:
                d = f32(f64(d) + 0.15)
            if d < 0 or d > f32(fs * f32(S)):
                continue
            b = int(np.floor(f32(f32(d / fs) / f32(2))))
            m = meas[i, b]
            th = f64(-2.0) * pi * f64(d) / f64(wavelength)
            pr, pim = f32(np.cos(th)), f32(np.sin(th))
            cr = f32(f32(m.real * pr) - f32(m.imag * pim))
            ci = f32(f32(m.real * pim) + f32(m.imag * pr))
            vec = (f32(f32(p[0] + f32(o[0] / f32(2))) - vx),
                   f32(f32(p[1] + f32(o[1] / f32(2))) - vy),
                   f32(f32(p[2] + f32(o[2] / f32(2))) - vz))
            w = f32(f32(cr * ir) + f32(ci * ii))
            w = f32(w / mag)
            for k in range(3):
                sn[k] = f32(sn[k] + f32(vec[k] * w))
        out[t] = sn
    return out
