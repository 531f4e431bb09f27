"""Drop-in for ``py_sn_est_gpu.normal_estimates_cuda`` (py_sn_est_gpu.py:382-424).

Same signature, defaults and return value as the reference: a float64
(NX, NY, NZ, 3) array of un-normalised votes, NaN where the voxel is not
above ``initial_filter_percent`` of the image maximum
(py_sn_est_gpu.py:326,374-376).  NumPy in, NumPy out.

* ``use_interpolated_processing=True`` (the reference's only mode), in
  antenna groups of 131072 rows whose votes are summed
  (py_sn_est_gpu.py:408-423): by default (``variant="fast"``) the vote from
  the moment M = sum_i vec_i c_i that ``mmn_bp_interp`` accumulates over the
  filtered voxels, n = Re(conj(I) M) / |I| with I the image passed in (the
  reference's per-row weight Re(c_i conj(I))/|I| folded over rows);
  ``variant="ref"`` runs the reference vote arithmetic
  (surface_normal_est.cu:64-149).
* ``use_interpolated_processing=False`` (the reference raises here,
  py_sn_est_gpu.py:404-405): the full-frequency vote, from the moment
  M = sum_i vec_i c_i that the recurrence kernel accumulates with the image
  (``variant="profile"``: the range-profile kernel), n = Re(conj(I) M) / |I|
  with I the image passed in.
"""
from __future__ import annotations

import math

import numpy as np

from . import ops
from .py_image_gpu import MAX_GROUP_ROWS, _device, _prep_inputs, _to_dev, host_bbox


def _filter(sum_image, initial_filter_percent):
    mag = np.abs(sum_image)
    return mag > (np.max(mag) * initial_filter_percent)


def _sn_est(sum_image, x_locs, y_locs, z_locs, antenna_locs, radar_data, rx_offset, slope,
            wavelength, bandwidth, num_samples, use_4_rx, is_ti_radar,
            initial_filter_percent, interpolated, K, variant="fast"):
    samples_per_meas = np.asarray(radar_data).shape[1]
    rx, meas, ant_rows = _prep_inputs(radar_data, rx_offset, antenna_locs, use_4_rx)
    filt = _filter(sum_image, initial_filter_percent)
    lin = np.flatnonzero(filt.ravel())  # C order == meshgrid/argwhere order of the reference
    nx, ny, nz = len(x_locs), len(y_locs), len(z_locs)
    all_normals = np.full((nx, ny, nz, 3), np.nan)
    if lin.size == 0:
        return all_normals
    dev = _device()
    xl, yl, zl = (_to_dev(a, np.float32, dev) for a in (x_locs, y_locs, z_locs))
    ant_t = _to_dev(np.asarray(ant_rows).reshape(-1, 3), np.float32, dev)
    rx_t = _to_dev(rx, np.float32, dev)
    meas_t = _to_dev(meas, np.complex64, dev)
    img_t = _to_dev(np.asarray(sum_image).ravel(), np.complex64, dev)
    vidx = _to_dev(lin, np.int32, dev)
    if interpolated:
        fft_spacing = np.float32(3e8 / (2 * bandwidth) * num_samples / samples_per_meas)
        if variant == "ref":
            normals = ops.sn_est(img_t, vidx, meas_t, ant_t, xl, yl, zl, rx_t,
                                 wavelength=np.float32(wavelength), fft_spacing=fft_spacing,
                                 is_ti=is_ti_radar)
        else:
            _, mom = ops.backproject_interp(meas_t, ant_t, xl, yl, zl, rx_t,
                                            wavelength=np.float32(wavelength),
                                            fft_spacing=fft_spacing, vox_idx=vidx,
                                            is_ti=is_ti_radar, moment=True)
            normals = ops.sn_from_moment(img_t, mom, vidx)
    else:
        f0 = ops.C0 / float(np.float32(wavelength))
        df = float(np.float32(slope))
        fwd = "profile" if variant == "profile" else "rec"
        bbox = host_bbox(x_locs, y_locs, z_locs) if fwd == "profile" else None
        _, mom = ops.backproject(meas_t, ant_t, xl, yl, zl, rx_t, f0=f0, df=df,
                                 variant=fwd, K=K, moment=True, vox_idx=vidx,
                                 is_ti=is_ti_radar, bbox=bbox)
        normals = ops.sn_from_moment(img_t, mom, vidx)
    all_normals[filt] = normals.cpu().numpy()
    return all_normals


def normal_estimates_cuda(sum_image, x_locs, y_locs, z_locs, antenna_locs, radar_data,
                          rx_offset, slope, wavelength, bandwidth, num_samples,
                          use_4_rx=False, is_ti_radar=False, initial_filter_percent=0.05,
                          use_interpolated_processing=True, *, K=64, variant="fast"):
    """Surface-normal votes on the GPU (see module docstring)."""
    if variant not in ("fast", "profile", "ref"):
        raise ValueError("variant must be 'fast', 'profile' or 'ref'")
    if not use_interpolated_processing:
        return _sn_est(sum_image, x_locs, y_locs, z_locs, antenna_locs, radar_data,
                       rx_offset, slope, wavelength, bandwidth, num_samples, use_4_rx,
                       is_ti_radar, initial_filter_percent, False, K, variant)
    num_rx = len(rx_offset)
    antenna_locs = np.asarray(antenna_locs)
    radar_data = np.asarray(radar_data)
    groups = math.ceil(antenna_locs.shape[0] * num_rx / MAX_GROUP_ROWS)
    normals = None
    for i in range(groups):
        lo, hi = i * MAX_GROUP_ROWS // num_rx, (i + 1) * MAX_GROUP_ROWS // num_rx
        sub = _sn_est(sum_image, x_locs, y_locs, z_locs, antenna_locs[lo:hi],
                      radar_data[lo:hi], rx_offset, slope, wavelength, bandwidth,
                      num_samples, use_4_rx, is_ti_radar, initial_filter_percent, True, K,
                      variant)
        normals = sub if normals is None else normals + sub
    return normals

