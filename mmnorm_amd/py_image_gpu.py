"""Drop-in for ``py_image_gpu.image_cuda`` (py_image_gpu.py:308-366).

Same signature, argument meaning, normalisation and return values as the
reference; NumPy arrays in and out.  The arithmetic runs in
libmmnorm_hip.so on the current ROCm device:

* full matched filter (``use_interpolated_processing=False``): by default the
  recurrence kernel (``variant="rec"``, FP32 Horner over samples with FP64
  re-anchoring every ``K`` samples); ``variant="profile"`` the same sum from
  per-row range-profile tables (mmn_bp_profile, ~13x faster, 1e-6 from
  FP64); ``variant="ref"`` runs the reference kernel's own arithmetic (FP64
  phase and cos/sin per sample) and ``variant="plain"`` the direct FP32
  sincosf kernel;
* interpolated range-bin branch: by default (and with ``"rec"``,
  ``"profile"``, ``"plain"``) ``mmn_bp_interp`` (the reference's FP32
  distances and bins, FP64-reduced phase with the hardware sine/cosine);
  ``variant="ref"`` runs the reference kernel's arithmetic.

Host steps mirror the reference: TX rows repeated per RX and measurements
reordered to (meas, rx, sample) when ``use_4_rx`` (py_image_gpu.py:228-234);
division by the TX count (py_image_gpu.py:302), a second time for the full
branch (py_image_gpu.py:330); antenna groups of 131072 rows for the
interpolated branch, each normalised by its own TX count before summing
(py_image_gpu.py:339-362); phase centres (p + (p + o)) / 2
(py_image_gpu.py:332-336).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib, ops

MAX_GROUP_ROWS = 4096 * 32  # py_image_gpu.py:339


def _prep_inputs(radar_data, rx_offset, antenna_locs, use_4_rx):
    """Layout of py_image_gpu.py:219-235 and py_sn_est_gpu.py:267-288: rows =
    meas x rx, samples last.

    The reference flattens the (transposed, when use_4_rx) measurements and
    its kernel reads row i at offset i * radar_data.shape[1]
    (imaging_gpu.cu:100,110), for len(antenna_locs) rows.  So without
    use_4_rx an (M, S) or (M, S, 1) array gives M rows of S samples, and
    with it an (M, S, R) array gives M*R rows (TX position repeated per RX).
    """
    radar_data = np.asarray(radar_data)
    antenna_locs = np.asarray(antenna_locs)
    if radar_data.ndim < 2:
        raise ValueError("radar_data must be (meas, samples) or (meas, samples, rx)")
    samples = radar_data.shape[1]
    if use_4_rx:
        radar_data = np.transpose(radar_data, (0, 2, 1))
        antenna_locs = np.repeat(antenna_locs, radar_data.shape[1], axis=0)
    rx_offset = np.asarray(rx_offset, dtype=np.float64).reshape(-1, 3)
    rows = len(antenna_locs)
    flat = radar_data.reshape(-1)
    if flat.size < rows * samples:
        raise ValueError(f"{flat.size} measurement values for {rows} rows of {samples} samples")
    meas = np.ascontiguousarray(flat[:rows * samples].reshape(rows, samples))
    return rx_offset, meas, antenna_locs


def _device():
    _lib.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def _to_dev(a, dtype, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)


def host_bbox(x_locs, y_locs, z_locs):
    """Bounding box of the FP32 voxel axes (what the device sees), from the
    host arrays: no read-back from the device."""
    ax = [np.asarray(a, dtype=np.float32).astype(np.float64) for a in (x_locs, y_locs, z_locs)]
    return [float(a.min()) for a in ax] + [float(a.max()) for a in ax]


def _run_image(x_locs, y_locs, z_locs, antenna_locs, radar_data, rx_offset, slope,
               wavelength, bandwidth, num_samples, use_4_rx, is_ti_radar, interpolated,
               variant, K):
    """py_image_gpu.py:237-300 on one device: returns (the kernel's image on the
    device, N, rows' TX positions); _reference_scaled applies :301-302."""
    norm_factor = np.asarray(antenna_locs).shape[0]
    samples_per_meas = np.asarray(radar_data).shape[1]
    rx, meas, ant_rows = _prep_inputs(radar_data, rx_offset, antenna_locs, use_4_rx)
    fft_spacing = np.float32(3e8 / (2 * bandwidth) * num_samples / samples_per_meas)
    dev = _device()
    xl, yl, zl = (_to_dev(a, np.float32, dev) for a in (x_locs, y_locs, z_locs))
    ant_t = _to_dev(np.asarray(ant_rows).reshape(-1, 3), np.float32, dev)
    rx_t = _to_dev(rx, np.float32, dev)
    meas_t = _to_dev(meas, np.complex64, dev)
    if meas_t.shape[0] != ant_t.shape[0]:
        raise ValueError(f"{meas_t.shape[0]} measurement rows for {ant_t.shape[0]} antenna rows")
    if variant == "ref":
        raw = ops.cuda_image(meas_t, ant_t, xl, yl, zl, rx_t, slope=np.float32(slope),
                             wavelength=np.float32(wavelength), fft_spacing=fft_spacing,
                             is_ti=is_ti_radar, interpolated=interpolated)
    elif interpolated:
        raw = ops.backproject_interp(meas_t, ant_t, xl, yl, zl, rx_t,
                                     wavelength=np.float32(wavelength),
                                     fft_spacing=fft_spacing, is_ti=is_ti_radar)
    else:
        f0 = ops.C0 / float(np.float32(wavelength))
        df = float(np.float32(slope))
        bbox = host_bbox(x_locs, y_locs, z_locs) if variant == "profile" else None
        raw = ops.backproject(meas_t, ant_t, xl, yl, zl, rx_t, f0=f0, df=df,
                              variant=variant, K=K, is_ti=is_ti_radar, bbox=bbox)
    return raw, norm_factor, ant_rows


_FLT_MIN = float(np.finfo(np.float32).tiny)


def _reference_scaled(raw, norm_factor, divisions):
    """The reference's host arithmetic on the kernel's output, bit for bit:
    ``p_xyz = p_xyz_r + 1j*p_xyz_i`` (complex64), then ``p_xyz /= norm_factor``
    ``divisions`` times (py_image_gpu.py:301-302, and :330 for the full
    branch).  Returns a complex64 NumPy array.

    NumPy divides a complex64 value by a real n as (re + im*rat)*scl and
    (im - re*rat)*scl with rat = 0/n and scl = float32(1)/float32(n); the
    ``+ 1j*`` step adds signed zeros.  When every component is finite and
    nonzero and stays a normal float through the divisions, both steps reduce
    to one float32 multiply by scl per component, which the device does
    before the copy to the host (one min/max reduction decides).  Any other
    image (zeros, NaN, infinities, values near the subnormal range) takes
    NumPy's own complex64 steps on the host."""
    n = int(norm_factor)
    scl = np.float32(1.0) / np.float32(n) if n >= 1 else None
    f = torch.view_as_real(raw)
    if scl is not None and f.numel():
        lo, hi = torch.aminmax(f.abs())
        lo, hi = float(lo), float(hi)
        if math.isfinite(hi) and lo * float(scl) ** divisions >= 2.0 * _FLT_MIN:
            for _ in range(divisions):
                f.mul_(float(scl))
            return raw.cpu().numpy()
    host = raw.cpu().numpy()
    p_xyz = host.real + 1j * host.imag
    p_xyz = p_xyz.astype(np.complex64, copy=False)
    for _ in range(divisions):
        p_xyz /= norm_factor
    return p_xyz


def _phase_centers(locs_rep, rx_offsets):
    rx_offsets = np.asarray(rx_offsets)
    rx_rep = np.tile(rx_offsets, reps=(len(locs_rep) // len(rx_offsets), 1))
    rx_locs = np.array(locs_rep, copy=True) + rx_rep
    return (locs_rep + rx_locs) / 2


def image_cuda(x_locs, y_locs, z_locs, antenna_locs, radar_data, rx_offsets, slope,
               wavelength, bandwidth, num_samples, use_4_rx=False, is_ti_radar=False,
               use_interpolated_processing=False, return_phase_centers=True, *,
               variant="rec", K=64):
    """SAR image on the GPU (see module docstring); returns (image, centers) or image."""
    if variant not in ("rec", "profile", "plain", "ref"):
        raise ValueError("variant must be 'rec', 'profile', 'plain' or 'ref'")
    if not use_interpolated_processing:
        raw, norm_factor, locs_rep = _run_image(
            x_locs, y_locs, z_locs, antenna_locs, radar_data, rx_offsets, slope,
            wavelength, bandwidth, num_samples, use_4_rx, is_ti_radar, False, variant, K)
        sum_image = _reference_scaled(raw, norm_factor, 2)
        center_locs = _phase_centers(locs_rep, rx_offsets)
    else:
        num_rx = len(rx_offsets)
        antenna_locs = np.asarray(antenna_locs)
        radar_data = np.asarray(radar_data)
        groups = math.ceil(antenna_locs.shape[0] * num_rx / MAX_GROUP_ROWS)
        sum_image, centers = None, []
        for i in range(groups):
            lo, hi = i * MAX_GROUP_ROWS // num_rx, (i + 1) * MAX_GROUP_ROWS // num_rx
            raw, norm_factor, locs_rep = _run_image(
                x_locs, y_locs, z_locs, antenna_locs[lo:hi], radar_data[lo:hi], rx_offsets,
                slope, wavelength, bandwidth, num_samples, use_4_rx, is_ti_radar, True,
                variant, K)
            img = _reference_scaled(raw, norm_factor, 1)
            if return_phase_centers:
                centers.append(_phase_centers(locs_rep, rx_offsets))
            sum_image = img if sum_image is None else sum_image + img
        center_locs = np.concatenate(centers, axis=0) if centers else None
    if return_phase_centers:
        return sum_image, center_locs
    return sum_image
