"""Torch-level API of the hot path (device tensors in, device tensors out).

Every function here calls libmmnorm_hip.so through the C ABI; there is no CPU
or PyTorch fallback.  Shapes:

* ``meas``      complex64 (rows, F)       measurement rows (row = tx * R + rx)
* ``ant``       float32  (rows, 3)        TX position of each row
* ``rx_off``    float32  (R, 3)           RX offset, row i uses rx_off[i % R]
* ``x/y/z_locs`` float32 (NX,), (NY,), (NZ,)  separable voxel grid
* ``vox_idx``   int32    (n,)             optional list of linear voxel indices

Frequencies: sample j is at f0 + j * df (Hz).  The reference's phase
-2 pi d / lambda - 2 pi d j slope / c (imaging_gpu.cu:112) is this with
f0 = c / lambda and df = slope.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

C0 = 2.99792458e8
# mmn_bp_forward's variant codes; "profile" is its own entry point, mmn_bp_profile
VARIANTS = {"plain": _lib.MMN_BP_PLAIN, "plain_fast": _lib.MMN_BP_PLAIN_FAST,
            "rec": _lib.MMN_BP_REC, "profile": None}


def _check(t, name, dtype, ndim=None, last=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on a ROCm device (got {t.device})")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if ndim is not None and t.dim() != ndim:
        raise ValueError(f"{name} must have {ndim} dims (got shape {tuple(t.shape)})")
    if last is not None and t.shape[-1] != last:
        raise ValueError(f"{name} must have last dim {last} (got shape {tuple(t.shape)})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def _check_index(idx, name, bound):
    """int32 index vector with every entry in [0, bound) (kernels index with it)."""
    _check(idx, name, torch.int32, 1)
    if idx.numel():
        lo, hi = torch.aminmax(idx)
        if int(lo) < 0 or int(hi) >= bound:
            raise ValueError(f"{name} entries must lie in [0, {bound})")
    return idx


def _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx):
    _lib.require_gpu()
    _check(meas, "meas", torch.complex64, 2)
    _check(ant, "ant", torch.float32, 2, 3)
    for n, t in (("x_locs", x_locs), ("y_locs", y_locs), ("z_locs", z_locs)):
        _check(t, n, torch.float32, 1)
        if t.numel() == 0:
            raise ValueError(f"{n} is empty")
    if rx_off is None:
        rx_off = torch.zeros((1, 3), dtype=torch.float32, device=meas.device)
    _check(rx_off, "rx_off", torch.float32, 2, 3)
    if ant.shape[0] != meas.shape[0]:
        raise ValueError(f"ant has {ant.shape[0]} rows, meas has {meas.shape[0]}")
    if meas.shape[0] == 0 or meas.shape[1] == 0:
        raise ValueError("meas must have at least one row and one sample")
    if rx_off.shape[0] == 0:
        raise ValueError("rx_off must hold at least one RX offset")
    nvox_grid = x_locs.numel() * y_locs.numel() * z_locs.numel()
    if vox_idx is not None:
        _check_index(vox_idx, "vox_idx", nvox_grid)
    if nvox_grid >= 2 ** 31:
        raise ValueError("voxel grid too large for int32 voxel indices")
    return rx_off, nvox_grid


def grid_bbox(x_locs, y_locs, z_locs):
    """(xmin, ymin, zmin, xmax, ymax, zmax) of the voxel axes, as host doubles."""
    lo_hi = [torch.aminmax(t) for t in (x_locs, y_locs, z_locs)]
    v = torch.stack([m for m, _ in lo_hi] + [M for _, M in lo_hi]).double().cpu()
    return [float(a) for a in v]


def backproject(meas, ant, x_locs, y_locs, z_locs, rx_off=None, *, f0, df,
                variant="rec", K=64, moment=False, vox_idx=None, is_ti=False, bbox=None):
    """Raw full-frequency sum  I_v = sum_ij meas_ij exp(-j 2 pi (f0 + j df) d_iv / c).

    Returns ``img`` complex64 (NX, NY, NZ) -- or (n,) with ``vox_idx`` -- and, if
    ``moment``, the vote moment complex64 (.., 3).  ``variant="profile"``
    evaluates the same sum from per-row range-profile tables (mmn_bp_profile);
    ``bbox`` (6 floats containing every voxel) saves it reading the axes'
    extent back from the device.
    """
    if variant == "profile":
        return _backproject_profile(meas, ant, x_locs, y_locs, z_locs, rx_off, f0=f0, df=df,
                                    moment=moment, vox_idx=vox_idx, is_ti=is_ti, bbox=bbox)
    rx_off, nvox_grid = _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx)
    if variant not in VARIANTS:
        raise ValueError(f"variant must be one of {sorted(VARIANTS)}")
    dev = meas.device
    n = vox_idx.numel() if vox_idx is not None else nvox_grid
    img = torch.empty(n, dtype=torch.complex64, device=dev)
    mom = torch.empty((n, 3), dtype=torch.complex64, device=dev) if moment else None
    if n:
        lib = _lib.load()
        ws_bytes = int(lib.mmn_bp_forward_workspace_size(n, meas.shape[0], int(moment)))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev) if ws_bytes else None
        _lib.call("mmn_bp_forward", img.data_ptr(), _lib.ptr(mom), x_locs.data_ptr(),
                  y_locs.data_ptr(), z_locs.data_ptr(), _lib.ptr(vox_idx), n,
                  ant.data_ptr(), meas.data_ptr(), rx_off.data_ptr(), float(f0), float(df),
                  x_locs.numel(), y_locs.numel(), z_locs.numel(), meas.shape[0],
                  rx_off.shape[0], meas.shape[1], int(bool(is_ti)), VARIANTS[variant],
                  int(K), _lib.ptr(ws), ws_bytes, _lib.stream_handle(dev))
    if vox_idx is None:
        img = img.view(x_locs.numel(), y_locs.numel(), z_locs.numel())
        if mom is not None:
            mom = mom.view(x_locs.numel(), y_locs.numel(), z_locs.numel(), 3)
    return (img, mom) if moment else img


def _backproject_profile(meas, ant, x_locs, y_locs, z_locs, rx_off, *, f0, df, moment,
                         vox_idx, is_ti, bbox):
    rx_off, nvox_grid = _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx)
    dev = meas.device
    n = vox_idx.numel() if vox_idx is not None else nvox_grid
    img = torch.empty(n, dtype=torch.complex64, device=dev)
    mom = torch.empty((n, 3), dtype=torch.complex64, device=dev) if moment else None
    if n:
        if bbox is None:
            bbox = grid_bbox(x_locs, y_locs, z_locs)
        if len(bbox) != 6:
            raise ValueError("bbox must hold 6 values")
        bb = (ctypes.c_double * 6)(*[float(v) for v in bbox])
        lib = _lib.load()
        ws_bytes = int(lib.mmn_bp_profile_workspace_size(n, meas.shape[0], meas.shape[1],
                                                         float(f0), float(df),
                                                         ctypes.addressof(bb),
                                                         int(moment)))
        if ws_bytes == 0:
            raise ValueError("mmn_bp_profile: bounding box not finite or too large for the "
                             "range-profile tables")
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        _lib.call("mmn_bp_profile", img.data_ptr(), _lib.ptr(mom), x_locs.data_ptr(),
                  y_locs.data_ptr(), z_locs.data_ptr(), _lib.ptr(vox_idx), n,
                  ant.data_ptr(), meas.data_ptr(), rx_off.data_ptr(), float(f0), float(df),
                  x_locs.numel(), y_locs.numel(), z_locs.numel(), meas.shape[0],
                  rx_off.shape[0], meas.shape[1], int(bool(is_ti)), ctypes.addressof(bb),
                  ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    if vox_idx is None:
        img = img.view(x_locs.numel(), y_locs.numel(), z_locs.numel())
        if mom is not None:
            mom = mom.view(x_locs.numel(), y_locs.numel(), z_locs.numel(), 3)
    return (img, mom) if moment else img


def cuda_image(meas, ant, x_locs, y_locs, z_locs, rx_off, *, slope, wavelength,
               fft_spacing=1.0, is_ti=False, interpolated=False):
    """The reference kernel's arithmetic (imaging_gpu.cu:48-122), raw sum (NX,NY,NZ)."""
    rx_off, nvox = _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, None)
    img = torch.empty(nvox, dtype=torch.complex64, device=meas.device)
    _lib.call("mmn_cuda_image", img.data_ptr(), x_locs.data_ptr(), y_locs.data_ptr(),
              z_locs.data_ptr(), ant.data_ptr(), meas.data_ptr(), rx_off.data_ptr(),
              float(slope), float(wavelength), float(fft_spacing), x_locs.numel(),
              y_locs.numel(), z_locs.numel(), meas.shape[0], rx_off.shape[0], meas.shape[1],
              int(bool(is_ti)), int(bool(interpolated)), _lib.stream_handle(meas.device))
    return img.view(x_locs.numel(), y_locs.numel(), z_locs.numel())


def backproject_interp(meas, ant, x_locs, y_locs, z_locs, rx_off=None, *, wavelength,
                       fft_spacing, vox_idx=None, is_ti=False, moment=False):
    """Interpolated (range-bin) backprojection, MI355X form: the reference's
    FP32 distance and bins, FP64-reduced phase with the hardware sine/cosine,
    and optionally the vote moment.  meas complex64 (rows, S) spectrum bins.
    Returns img (NX,NY,NZ) (or (n,) for vox_idx) and, with moment=True, also
    the moment complex64 (n, 3)."""
    rx_off, nvox = _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx)
    n = vox_idx.numel() if vox_idx is not None else nvox
    dev = meas.device
    img = torch.empty(n, dtype=torch.complex64, device=dev)
    mom = torch.empty((n, 3), dtype=torch.complex64, device=dev) if moment else None
    _lib.call("mmn_bp_interp", img.data_ptr(), _lib.ptr(mom), x_locs.data_ptr(),
              y_locs.data_ptr(), z_locs.data_ptr(), _lib.ptr(vox_idx), n, ant.data_ptr(),
              meas.data_ptr(), rx_off.data_ptr(), float(wavelength), float(fft_spacing),
              x_locs.numel(), y_locs.numel(), z_locs.numel(), meas.shape[0], rx_off.shape[0],
              meas.shape[1], int(bool(is_ti)), _lib.stream_handle(dev))
    if vox_idx is None:
        img = img.view(x_locs.numel(), y_locs.numel(), z_locs.numel())
    return (img, mom) if moment else img


def sn_est(img, vox_idx, meas, ant, x_locs, y_locs, z_locs, rx_off, *, wavelength,
           fft_spacing, is_ti=False):
    """The reference vote (surface_normal_est.cu:64-149) over vox_idx, float32 (n, 3)."""
    rx_off, nvox = _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx)
    _check(img, "img", torch.complex64)
    if img.numel() != nvox:
        raise ValueError("img must cover the whole voxel grid")
    n = vox_idx.numel()
    out = torch.empty((n, 3), dtype=torch.float32, device=meas.device)
    _lib.call("mmn_cuda_sn_est", out.data_ptr(), vox_idx.data_ptr(), n, img.data_ptr(),
              x_locs.data_ptr(), y_locs.data_ptr(), z_locs.data_ptr(), ant.data_ptr(),
              meas.data_ptr(), rx_off.data_ptr(), 0.0, float(wavelength),
              float(fft_spacing), x_locs.numel(), y_locs.numel(), z_locs.numel(),
              meas.shape[0], rx_off.shape[0], meas.shape[1], int(bool(is_ti)),
              _lib.stream_handle(meas.device))
    return out


def sn_from_moment(img, moment, vox_idx):
    """Re(conj(I) M) / |I| with I = img.flatten()[vox_idx]; float32 (n, 3)."""
    _lib.require_gpu()
    _check(img, "img", torch.complex64)
    _check(moment, "moment", torch.complex64, 2, 3)
    _check_index(vox_idx, "vox_idx", img.numel())
    if moment.shape[0] != vox_idx.numel():
        raise ValueError("moment rows must match vox_idx")
    n = vox_idx.numel()
    out = torch.empty((n, 3), dtype=torch.float32, device=img.device)
    _lib.call("mmn_sn_from_moment", out.data_ptr(), img.data_ptr(), moment.data_ptr(),
              vox_idx.data_ptr(), n, _lib.stream_handle(img.device))
    return out


def backproject_backward(g, meas, ant, x_locs, y_locs, z_locs, rx_off=None, *, f0, df,
                         scale=1.0, K=None, vox_idx=None, is_ti=False, need_meas=True,
                         need_ant=True, variant="rec", bbox=None):
    """Adjoint of ``scale * backproject``: (dL/dmeas complex64 (rows,F), dL/dant float32 (rows,3)).

    ``K`` is the recurrence backward's sample block: 16 or 32, or 64 for the
    position gradient alone; by default 64 when only dL/dant is asked for
    (no adjoint accumulators, so the longer block pays) and 32 otherwise.

    ``variant="rec"`` recomputes every (row, sample, voxel) phase
    (mmn_bp_backward); ``variant="profile"`` evaluates the same gradients by
    voxel groups with short series in the distance (mmn_bp_profile_backward;
    it synchronises the stream once), ``bbox`` as for the profile forward.
    """
    if variant not in ("rec", "profile"):
        raise ValueError('variant must be "rec" or "profile"')
    rx_off, nvox_grid = _geometry(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx)
    n = vox_idx.numel() if vox_idx is not None else nvox_grid
    g = g.reshape(-1)
    _check(g, "grad", torch.complex64, 1)
    if g.numel() != n:
        raise ValueError(f"grad has {g.numel()} entries, expected {n}")
    dev = meas.device
    A, F = meas.shape
    if K is None:
        K = 64 if (need_ant and not need_meas) else 32
    gm = torch.empty((A, F), dtype=torch.complex64, device=dev) if need_meas else None
    ga = torch.empty((A, 3), dtype=torch.float32, device=dev) if need_ant else None
    lib = _lib.load()
    if variant == "profile":
        if bbox is None:
            bbox = grid_bbox(x_locs, y_locs, z_locs)
        if len(bbox) != 6:
            raise ValueError("bbox must hold 6 values")
        bb = (ctypes.c_double * 6)(*[float(v) for v in bbox])
        ws_bytes = int(lib.mmn_bp_profile_backward_workspace_size(
            max(n, 1), A, F, float(f0), float(df), ctypes.addressof(bb)))
        if ws_bytes == 0:
            raise ValueError("mmn_bp_profile_backward: sizes must be positive")
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        _lib.call("mmn_bp_profile_backward", _lib.ptr(gm), _lib.ptr(ga), g.data_ptr(),
                  float(scale), x_locs.data_ptr(), y_locs.data_ptr(), z_locs.data_ptr(),
                  _lib.ptr(vox_idx), n, ant.data_ptr(), meas.data_ptr(), rx_off.data_ptr(),
                  float(f0), float(df), x_locs.numel(), y_locs.numel(), z_locs.numel(), A,
                  rx_off.shape[0], F, int(bool(is_ti)), ctypes.addressof(bb), ws.data_ptr(),
                  ws_bytes, _lib.stream_handle(dev))
        return gm, ga
    ws_bytes = int(lib.mmn_bp_backward_workspace_size(A, F, n))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    _lib.call("mmn_bp_backward", _lib.ptr(gm), _lib.ptr(ga), g.data_ptr(), float(scale),
              x_locs.data_ptr(), y_locs.data_ptr(), z_locs.data_ptr(), _lib.ptr(vox_idx), n,
              ant.data_ptr(), meas.data_ptr(), rx_off.data_ptr(), float(f0), float(df),
              x_locs.numel(), y_locs.numel(), z_locs.numel(), A, rx_off.shape[0], F,
              int(bool(is_ti)), int(K), ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    return gm, ga


class Backproject(torch.autograd.Function):
    """I = scale * backproject(meas, ant, ...), differentiable in meas and ant.

    The backward kernel recomputes the phases (no rows x samples x voxels
    tensor is stored); gradients follow PyTorch's complex convention.
    """

    @staticmethod
    def forward(ctx, meas, ant, x_locs, y_locs, z_locs, rx_off, f0, df, scale, K, vox_idx,
                is_ti, variant, bbox):
        if variant == "profile" and bbox is None:
            bbox = grid_bbox(x_locs, y_locs, z_locs)
        img = backproject(meas, ant, x_locs, y_locs, z_locs, rx_off, f0=f0, df=df,
                          variant=variant, K=K, vox_idx=vox_idx, is_ti=is_ti, bbox=bbox)
        ctx.save_for_backward(meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx)
        # the backward's sample block: its default for the outputs asked for
        # (64 for the position gradient alone, else 32), 16 after a short forward block
        ctx.args = (f0, df, scale, None if K >= 32 else 16, is_ti, variant, bbox)
        return img * scale

    @staticmethod
    def backward(ctx, grad):
        meas, ant, x_locs, y_locs, z_locs, rx_off, vox_idx = ctx.saved_tensors
        f0, df, scale, K, is_ti, variant, bbox = ctx.args
        need_m, need_a = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_m or need_a):
            return (None,) * 14
        gm, ga = backproject_backward(grad.contiguous(), meas, ant, x_locs, y_locs, z_locs,
                                      rx_off, f0=f0, df=df, scale=scale, K=K,
                                      vox_idx=vox_idx, is_ti=is_ti, need_meas=need_m,
                                      need_ant=need_a, variant=variant, bbox=bbox)
        return (gm, ga) + (None,) * 12


def backproject_ad(meas, ant, x_locs, y_locs, z_locs, rx_off=None, *, f0, df, scale=1.0,
                   K=64, vox_idx=None, is_ti=False, variant="rec", bbox=None):
    """Differentiable ``scale * backproject``: the recurrence forward and its
    backward (``variant="rec"``), or the range-profile forward and the
    grouped-series backward (``variant="profile"``)."""
    if variant not in ("rec", "profile"):
        raise ValueError('variant must be "rec" or "profile"')
    if rx_off is None:
        rx_off = torch.zeros((1, 3), dtype=torch.float32, device=meas.device)
    return Backproject.apply(meas, ant, x_locs, y_locs, z_locs, rx_off, f0, df, scale, K,
                             vox_idx, is_ti, variant, bbox)


def rsdf_bfs(normals, labels, seeds, dx):
    """FIFO-exact RSDF breadth-first search on the GPU.

    normals float64 (NX,NY,NZ,3), labels int32 (NX,NY,NZ), seeds int32 (n,)
    linear start voxels in component order.  Returns (rsdf float64, hops int32),
    both (NX,NY,NZ).
    """
    _lib.require_gpu()
    _check(normals, "normals", torch.float64, 4, 3)
    _check(labels, "labels", torch.int32, 3)
    _check(seeds, "seeds", torch.int32, 1)
    if tuple(normals.shape[:3]) != tuple(labels.shape):
        raise ValueError("normals and labels must cover the same grid")
    nx, ny, nz = labels.shape
    V = nx * ny * nz
    if seeds.numel():
        lo, hi = int(seeds.min()), int(seeds.max())
        if lo < 0 or hi >= V:
            raise ValueError("seed index outside the grid")
    dev = labels.device
    rsdf = torch.empty((nx, ny, nz), dtype=torch.float64, device=dev)
    hops = torch.empty((nx, ny, nz), dtype=torch.int32, device=dev)
    lib = _lib.load()
    ws_bytes = int(lib.mmn_rsdf_bfs_workspace_size(nx, ny, nz))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    _lib.call("mmn_rsdf_bfs", rsdf.data_ptr(), hops.data_ptr(), normals.data_ptr(),
              labels.data_ptr(), seeds.data_ptr(), seeds.numel(), float(dx), nx, ny, nz,
              ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    return rsdf, hops


def li_threshold(img, exponent, max_iter=1000, below=False):
    """Li threshold of |img|**exponent on the GPU.

    img complex64 (any shape).  Returns (threshold, iterations, converged):
    the threshold in the units of |img|**exponent, the number of Li
    iterations run (at most ``max_iter``) and whether the stopping rule was
    met; with ``below=True`` also the bool tensor |img|**exponent < threshold
    (shaped like img).  Synchronises the stream once per 16 iterations.
    """
    _lib.require_gpu()
    _check(img, "img", torch.complex64)
    n = img.numel()
    if n == 0:
        raise ValueError("img is empty")
    import ctypes
    dev = img.device
    lib = _lib.load()
    ws_bytes = int(lib.mmn_li_threshold_workspace_size(n))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    thr, it, conv = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
    mask = torch.empty(img.shape, dtype=torch.uint8, device=dev) if below else None
    _lib.call("mmn_li_threshold", ctypes.addressof(thr), ctypes.addressof(it),
              ctypes.addressof(conv), _lib.ptr(mask), img.data_ptr(), n, float(exponent),
              int(max_iter), ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    if below:
        return thr.value, it.value, bool(conv.value), mask.view(torch.bool)
    return thr.value, it.value, bool(conv.value)


def ccl6(mask):
    """6-connected components of a 3-D mask on the GPU.

    mask bool or uint8 (NX,NY,NZ).  Returns (labels int32 (NX,NY,NZ), n):
    0 on background, 1..n numbered in the order a C-order scan first meets
    each component.
    """
    _lib.require_gpu()
    if not isinstance(mask, torch.Tensor):
        raise TypeError("mask must be a torch.Tensor")
    if mask.dtype == torch.bool:
        mask = mask.view(torch.uint8)
    _check(mask, "mask", torch.uint8, 3)
    import ctypes
    nx, ny, nz = mask.shape
    dev = mask.device
    labels = torch.empty((nx, ny, nz), dtype=torch.int32, device=dev)
    lib = _lib.load()
    ws_bytes = int(lib.mmn_ccl6_workspace_size(nx, ny, nz))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    n = ctypes.c_int32()
    _lib.call("mmn_ccl6", labels.data_ptr(), ctypes.addressof(n), mask.data_ptr(), nx, ny, nz,
              ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    return labels, int(n.value)


def label_sizes(labels, n_labels):
    """int64 (n_labels + 1,): entries of ``labels`` equal to l for l = 1..n_labels;
    index 0 counts all other entries."""
    _lib.require_gpu()
    _check(labels, "labels", torch.int32)
    sizes = torch.empty(n_labels + 1, dtype=torch.int64, device=labels.device)
    _lib.call("mmn_label_sizes", sizes.data_ptr(), labels.data_ptr(), labels.numel(),
              int(n_labels), _lib.stream_handle(labels.device))
    return sizes


def _f32s(vals, n):
    import ctypes
    vals = [float(v) for v in vals]
    if len(vals) != n:
        raise ValueError(f"expected {n} values, got {len(vals)}")
    return (ctypes.c_float * n)(*vals)


def ray_matching(grid, width, height, width_fov, height_fov, bounding_box, eye, rx_dir):
    """Sphere-trace width x height rays from ``eye`` through the SDF ``grid``.

    grid float32 (NX,NY,NZ); bounding_box (min x, y, z, max x, y, z); eye and
    rx_dir 3 values.  Returns (intersection_pos, voxel_position, directions),
    each float32 (width, height, 3) in the reference's pixel order; -1 where a
    ray has no hit.
    """
    import ctypes
    _lib.require_gpu()
    _check(grid, "grid", torch.float32, 3)
    dev = grid.device
    n = int(width) * int(height)
    hit, vox, dirs = (torch.empty((width, height, 3), dtype=torch.float32, device=dev)
                      for _ in range(3))
    res = (ctypes.c_int * 3)(*grid.shape)
    _lib.call("mmn_ray_matching", hit.data_ptr(), vox.data_ptr(), dirs.data_ptr(),
              grid.data_ptr(), int(width), int(height), float(width_fov), float(height_fov),
              _f32s(bounding_box, 6), res, _f32s(eye, 3), _f32s(rx_dir, 3),
              _lib.stream_handle(dev))
    assert n == hit.shape[0] * hit.shape[1]
    return hit, vox, dirs


def simulate_channel(grid, tau, width, height, width_fov, height_fov, bounding_box, eye,
                     rx_dir, rx_loc, voxel_size, f0, df, samples, is_ti=False):
    """Channel (complex64, samples) that one antenna receives from the isosurface
    grid - tau: rays marched, hits refined, sum of exp(+j 2 pi (f0 + j df) d / c)
    over hits with d the FP64 round trip (+0.15 m if is_ti)."""
    import ctypes
    _lib.require_gpu()
    _check(grid, "grid", torch.float32, 3)
    dev = grid.device
    ch = torch.empty(int(samples), dtype=torch.complex64, device=dev)
    lib = _lib.load()
    ws_bytes = int(lib.mmn_simulate_channel_workspace_size(int(width), int(height),
                                                           int(samples)))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    res = (ctypes.c_int * 3)(*grid.shape)
    rx = (ctypes.c_double * 3)(*[float(v) for v in rx_loc])
    _lib.call("mmn_simulate_channel", ch.data_ptr(), grid.data_ptr(), float(tau), int(width),
              int(height), float(width_fov), float(height_fov), _f32s(bounding_box, 6), res,
              _f32s(eye, 3), _f32s(rx_dir, 3), rx, float(voxel_size), float(f0), float(df),
              int(samples), int(bool(is_ti)), ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    return ch


MAX_BATCH_RAYS = 1 << 27   # rays of one mmn_simulate_channels launch (1 GiB of distances)


def simulate_channels(grid, comp, taus, eyes, rx_dirs, rx_locs, width, height, width_fov,
                      height_fov, bounding_box, voxel_size, f0, df, samples, is_ti=False):
    """Channels of many simulations at once (complex64, (n_sim, samples)):
    simulation s is simulate_channel on component ``grid[comp[s]]`` with
    offset ``taus[s]``, antenna ``eyes[s]``, direction ``rx_dirs[s]`` and
    receiver ``rx_locs[s]``, bit for bit.  ``grid``: float32 (n_comp, NX, NY,
    NZ) on the device; the per-simulation values are host arrays.  Launches
    hold at most 65535 simulations and MAX_BATCH_RAYS rays."""
    _lib.require_gpu()
    _check(grid, "grid", torch.float32, 4)
    dev = grid.device
    comp = np.ascontiguousarray(comp, dtype=np.int32).reshape(-1)
    n = len(comp)
    taus = np.ascontiguousarray(taus, dtype=np.float32).reshape(n)
    eyes = np.ascontiguousarray(eyes, dtype=np.float32).reshape(n, 3)
    rx_dirs = np.ascontiguousarray(rx_dirs, dtype=np.float32).reshape(n, 3)
    rx_locs = np.ascontiguousarray(rx_locs, dtype=np.float64).reshape(n, 3)
    if n and (comp.min() < 0 or comp.max() >= grid.shape[0]):
        raise ValueError(f"comp must lie in [0, {grid.shape[0]})")
    out = torch.empty((n, int(samples)), dtype=torch.complex64, device=dev)
    lib = _lib.load()
    res = (ctypes.c_int * 3)(*grid.shape[1:])
    bb = _f32s(bounding_box, 6)
    for lo, hi in _batch_groups(n, int(width) * int(height)):
        ws_bytes = int(lib.mmn_simulate_channels_workspace_size(int(width), int(height),
                                                                int(samples), hi - lo))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.call("mmn_simulate_channels", out[lo:hi].data_ptr(), grid.data_ptr(),
                  int(grid.shape[0]), comp[lo:hi].ctypes.data, taus[lo:hi].ctypes.data,
                  eyes[lo:hi].ctypes.data, rx_dirs[lo:hi].ctypes.data,
                  rx_locs[lo:hi].ctypes.data, hi - lo, int(width), int(height),
                  float(width_fov), float(height_fov), bb, res, float(voxel_size), float(f0),
                  float(df), int(samples), int(bool(is_ti)), ws.data_ptr(), ws_bytes,
                  _lib.stream_handle(dev))
    return out


class RayPrefix:
    """The offset-independent march of a set of simulations (mmn_ray_prefix):
    per launch group of simulations, the device buffer and the host live-ray
    counts, with the simulations' components, antennas and directions and the
    geometry they were marched with."""

    def __init__(self, grid, comp, eyes, rx_dirs, width, height, width_fov, height_fov,
                 bounding_box, tau_lo, groups):
        self.grid, self.comp, self.eyes, self.rx_dirs = grid, comp, eyes, rx_dirs
        self.width, self.height = int(width), int(height)
        self.width_fov, self.height_fov = float(width_fov), float(height_fov)
        self.bounding_box = [float(v) for v in bounding_box]
        self.tau_lo = float(np.float32(tau_lo))
        self.groups = groups  # [(lo, hi, buffer, live_counts)]

    @property
    def live_rays(self):
        return int(sum(int(g[3].sum()) for g in self.groups))


def _batch_groups(n, rays):
    per = max(1, min(65535, MAX_BATCH_RAYS // max(rays, 1)))
    return [(lo, min(n, lo + per)) for lo in range(0, n, per)]


def ray_prefix(grid, comp, eyes, rx_dirs, width, height, width_fov, height_fov, bounding_box,
               tau_lo):
    """March the rays of simulations (comp[s], eyes[s], rx_dirs[s]) as far as
    no offset >= ``tau_lo`` changes them (mmn_ray_prefix); the result feeds
    simulate_channels_prefixed for any such offsets."""
    _lib.require_gpu()
    _check(grid, "grid", torch.float32, 4)
    dev = grid.device
    comp = np.ascontiguousarray(comp, dtype=np.int32).reshape(-1)
    n = len(comp)
    eyes = np.ascontiguousarray(eyes, dtype=np.float32).reshape(n, 3)
    rx_dirs = np.ascontiguousarray(rx_dirs, dtype=np.float32).reshape(n, 3)
    if n and (comp.min() < 0 or comp.max() >= grid.shape[0]):
        raise ValueError(f"comp must lie in [0, {grid.shape[0]})")
    if not np.isfinite(tau_lo):
        raise ValueError("tau_lo must be finite")
    lib = _lib.load()
    res = (ctypes.c_int * 3)(*grid.shape[1:])
    bb = _f32s(bounding_box, 6)
    groups = []
    for lo, hi in _batch_groups(n, int(width) * int(height)):
        nb = int(lib.mmn_ray_prefix_size(int(width), int(height), hi - lo))
        buf = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
        live = np.zeros(hi - lo, dtype=np.int32)
        _lib.call("mmn_ray_prefix", buf.data_ptr(), nb, live.ctypes.data, grid.data_ptr(),
                  int(grid.shape[0]), comp[lo:hi].ctypes.data, eyes[lo:hi].ctypes.data,
                  rx_dirs[lo:hi].ctypes.data, hi - lo, int(width), int(height),
                  float(width_fov), float(height_fov), bb, res, float(np.float32(tau_lo)),
                  _lib.stream_handle(dev))
        groups.append((lo, hi, buf, live))
    return RayPrefix(grid, comp, eyes, rx_dirs, width, height, width_fov, height_fov,
                     bounding_box, tau_lo, groups)


def simulate_channels_prefixed(prefix, taus, rx_locs, voxel_size, f0, df, samples,
                               is_ti=False):
    """simulate_channels for the simulations of ``prefix`` with offsets
    ``taus[s]`` (each >= prefix.tau_lo) and receivers ``rx_locs[s]``, bit for
    bit, marching and summing only the rays the prefix left live."""
    _lib.require_gpu()
    n = len(prefix.comp)
    taus = np.ascontiguousarray(taus, dtype=np.float32).reshape(n)
    rx_locs = np.ascontiguousarray(rx_locs, dtype=np.float64).reshape(n, 3)
    if n and taus.min() < np.float32(prefix.tau_lo):
        raise ValueError(f"taus must be >= the prefix's tau_lo ({prefix.tau_lo})")
    grid = prefix.grid
    dev = grid.device
    out = torch.empty((n, int(samples)), dtype=torch.complex64, device=dev)
    lib = _lib.load()
    res = (ctypes.c_int * 3)(*grid.shape[1:])
    bb = _f32s(prefix.bounding_box, 6)
    W, H = prefix.width, prefix.height
    for lo, hi, buf, live in prefix.groups:
        ws_bytes = int(lib.mmn_simulate_channels_prefixed_workspace_size(
            W, H, int(samples), hi - lo, int(live.max()) if len(live) else 0))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        _lib.call("mmn_simulate_channels_prefixed", out[lo:hi].data_ptr(), buf.data_ptr(),
                  buf.numel(), live.ctypes.data, grid.data_ptr(), int(grid.shape[0]),
                  prefix.comp[lo:hi].ctypes.data, taus[lo:hi].ctypes.data,
                  prefix.eyes[lo:hi].ctypes.data, prefix.rx_dirs[lo:hi].ctypes.data,
                  rx_locs[lo:hi].ctypes.data, hi - lo, W, H, prefix.width_fov,
                  prefix.height_fov, bb, res, float(voxel_size), float(f0), float(df),
                  int(samples), int(bool(is_ti)), prefix.tau_lo, ws.data_ptr(), ws_bytes,
                  _lib.stream_handle(dev))
    return out


def channel_loss(sim, target, n_fft, bin_lo, bin_hi, bin_hi_target=None):
    """Range-profile L1 loss of each (sim, target) channel pair: complex64 (B, F)
    each -> float32 (B,)."""
    _lib.require_gpu()
    sim = sim.reshape(-1, sim.shape[-1]).contiguous()
    target = target.reshape(-1, target.shape[-1]).contiguous()
    _check(sim, "sim", torch.complex64, 2)
    _check(target, "target", torch.complex64, 2)
    if sim.shape != target.shape:
        raise ValueError("sim and target must have the same shape")
    B, F = sim.shape
    out = torch.empty(B, dtype=torch.float32, device=sim.device)
    hi_t = bin_hi if bin_hi_target is None else int(bin_hi_target)
    _lib.call("mmn_channel_loss", out.data_ptr(), sim.data_ptr(), target.data_ptr(), B, F,
              int(n_fft), int(bin_lo), int(bin_hi), hi_t, _lib.stream_handle(sim.device))
    return out
