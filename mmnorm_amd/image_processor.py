"""The caller of the hot path, ``ImageProcessor.generate_mmwave_surface_normals``
(image_processor.py:296-398), without its dataset loader: radar-frame
antenna positions, the 1 mm imaging grid, the RX offsets and radar
constants of the 77 GHz rig, the optional range FFT of the interpolated
branch (image_processor.py:22-35), the image and the vote, and the unit
normals -- with every intermediate kept on the GPU (the reference passes
NumPy arrays between the two calls and to the FFT).

``interpolate_measurements`` is the reference's zero-padded FFT along the
samples, done by the device FFT on complex64 (the reference: NumPy,
complex128, then narrowed to float32 by the image kernel).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, ops
from .py_image_gpu import MAX_GROUP_ROWS

C0 = 2.99792458e8
LOCATION0 = (0.42, -0.23, 0.23)          # image_processor.py:317
GRID = dict(min_x=-0.1, max_x=0.45, min_y=-0.05, max_y=0.55, min_z=-0.5, max_z=-0.25,
            step=0.001)                  # image_processor.py:327-336
SLOPE_REAL = 58.6116 * 1e6 * 1e6 / 10e6  # image_processor.py:358
TX_OFFSET = 0.005                        # image_processor.py:359


def _dev():
    _lib.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def interpolate_measurements(radar_data, FFT_SIZE=4096, as_tensor=False):
    """Zero-pad (N, S, R) measurements to FFT_SIZE samples and FFT along them
    (image_processor.py:22-35), on the GPU.  Returns complex64 (N, FFT_SIZE, R):
    NumPy by default, the device tensor with as_tensor=True."""
    x = radar_data if torch.is_tensor(radar_data) else torch.from_numpy(
        np.ascontiguousarray(radar_data, dtype=np.complex64))
    x = x.to(_dev(), torch.complex64)
    if x.shape[1] > FFT_SIZE:
        raise ValueError("more samples than FFT_SIZE")
    out = torch.fft.fft(x, n=FFT_SIZE, dim=1)
    return out if as_tensor else out.cpu().numpy()


def rx_offsets_for(aperture_type, wavelength, is_sim=False):
    """RX offsets of the rig (image_processor.py:357-376)."""
    if is_sim:
        return np.array([[0.0, 0.0, 0.0]])
    r = wavelength / 2
    steps = [TX_OFFSET + r * 3, TX_OFFSET + r * 2, TX_OFFSET + r, TX_OFFSET]
    if aperture_type == "original":
        return np.array([[0, -s, 0] for s in steps], dtype=np.float64)
    if aperture_type == "almost-square":
        return np.array([[-s, 0, 0] for s in steps], dtype=np.float64)
    raise ValueError(f"unknown aperture type {aperture_type!r}")


def imaging_grid(grid=GRID):
    g = grid
    return (np.arange(g["min_x"], g["max_x"], g["step"]), np.arange(g["min_y"], g["max_y"], g["step"]),
            np.arange(g["min_z"], g["max_z"], g["step"]))


def generate_mmwave_surface_normals(poses, radar_data, aperture_type, min_f, num_samples,
                                    bandwidth, slope=None, is_sim=False,
                                    use_interpolated_processing=True, location0=LOCATION0,
                                    grid=GRID, K=64):
    """The SAR image, the grid axes, unit surface normals (NaN outside the
    5 % filter), the rows' phase centres and the clipped measurements, as
    image_processor.py:296-398 returns them, from the correlated poses
    (N, >= 3) and measurements (N, >= num_samples, >= R) the loader would
    give it.  Radar constants come in as arguments (the reference reads
    them from params.json); poses are robot-frame (the motion-capture frame
    of image_processor.py:341 is not supported)."""
    dev = _dev()
    locs = np.asarray(poses, dtype=np.float64)[:, :3] - np.asarray(location0)
    max_rx = 1 if is_sim else 4
    meas_np = np.asarray(radar_data)[:, :num_samples, :max_rx]
    wavelength = C0 / min_f
    if is_sim:  # image_processor.py:373-376
        slope = bandwidth / (num_samples - 1) if slope is None else slope
    else:
        slope = SLOPE_REAL
    rxo = rx_offsets_for(aperture_type, wavelength, is_sim)
    xl, yl, zl = imaging_grid(grid)
    R = meas_np.shape[2]
    N = meas_np.shape[0]
    meas_t = torch.from_numpy(np.ascontiguousarray(meas_np, dtype=np.complex64)).to(dev)
    if use_interpolated_processing:
        spec = interpolate_measurements(meas_t, as_tensor=True)        # (N, 4096, R)
        samples_per_meas = spec.shape[1]
    else:
        spec = meas_t
        samples_per_meas = num_samples
    # rows = meas x rx, samples last (py_image_gpu.py:228-234); TX repeated per RX
    rows = spec.permute(0, 2, 1).reshape(N * R, samples_per_meas).contiguous()
    ant_rows = np.repeat(locs, R, axis=0)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    ant_t, rx_t = T(ant_rows, np.float32), T(rxo, np.float32)
    xt, yt, zt = T(xl, np.float32), T(yl, np.float32), T(zl, np.float32)
    lam32 = np.float32(wavelength)
    is_ti = not is_sim
    if use_interpolated_processing:
        fs = np.float32(3e8 / (2 * bandwidth) * num_samples / samples_per_meas)
        # groups of 131072 rows, each divided by its own TX count, summed
        # (py_image_gpu.py:339-362)
        per = MAX_GROUP_ROWS // R
        img = None
        for lo in range(0, N, per):
            hi = min(N, lo + per)
            part = ops.backproject_interp(rows[lo * R:hi * R], ant_t[lo * R:hi * R], xt, yt, zt,
                                          rx_t, wavelength=lam32, fft_spacing=fs,
                                          is_ti=is_ti) / (hi - lo)
            img = part if img is None else img + part
    else:
        img = ops.backproject(rows, ant_t, xt, yt, zt, rx_t, f0=C0 / float(lam32),
                              df=float(np.float32(slope)), K=K, is_ti=is_ti)
        img = img / N / N                      # py_image_gpu.py:302,330
    # normals on the 5 % filter (py_sn_est_gpu.py:326), the moment over those voxels
    mag = img.abs().reshape(-1)
    vidx = torch.nonzero(mag > 0.05 * mag.max()).reshape(-1).to(torch.int32)
    if use_interpolated_processing:
        _, mom = ops.backproject_interp(rows, ant_t, xt, yt, zt, rx_t, wavelength=lam32,
                                        fft_spacing=fs, vox_idx=vidx, is_ti=is_ti, moment=True)
    else:
        _, mom = ops.backproject(rows, ant_t, xt, yt, zt, rx_t, f0=C0 / float(lam32),
                                 df=float(np.float32(slope)), K=K, vox_idx=vidx, is_ti=is_ti,
                                 moment=True)
    nv = ops.sn_from_moment(img.reshape(-1).contiguous(), mom, vidx).double()
    nv = nv / torch.linalg.norm(nv, dim=1, keepdim=True)
    normals = torch.full((img.numel(), 3), float("nan"), dtype=torch.float64, device=dev)
    normals[vidx.long()] = nv
    rx_rep = np.tile(rxo, (N, 1))
    centers = (ant_rows + (ant_rows + rx_rep)) / 2
    return (img.cpu().numpy(), (xl, yl, zl), normals.reshape(*img.shape, 3).cpu().numpy(),
            centers, meas_np)


def correlate_locs_and_meas(robot_data, radar_data, loc_times, locs, num_frames, sweep_time,
                            periodicity, samples_per_chirp, num_chirps, min_f,
                            is_optitrack=False, num_lanes=4):
    """Pose of every radar frame by linear interpolation of the robot
    trajectory (image_processor.py:74-154), 77 GHz layout.

    robot_data: {'all_radar_filenames', 'all_radar_created_ts'}; radar_data:
    {int(filename): (samples, num_lanes) complex} as read_adc_bin gives;
    loc_times (T,), locs (T, D) the trajectory.  Frame j of a file starts at
    created_ts + j (sweep_time + periodicity); frames outside the trajectory's
    time span, frames within 3/16 wavelength of an earlier frame of the same
    file, and frames with too few samples are dropped -- and, as in the
    reference, once a file's first frame is dropped the duplicate test fails
    for its later frames, which are dropped too.  Returns {'radar_data': list
    of (samples_per_chirp, num_lanes) arrays, 'poses': (n, D)}.
    """
    loc_times = np.asarray(loc_times, dtype=np.float64)
    locs = np.asarray(locs, dtype=np.float64)
    thr = C0 / min_f * (3 / 16)
    t_lo, t_hi = loc_times.min(), loc_times.max()
    all_locs, all_data = [], []
    for i, fname in enumerate(robot_data["all_radar_filenames"]):
        ts = robot_data["all_radar_created_ts"][i]
        if ts == -1 or int(fname) not in radar_data:
            continue
        data = radar_data[int(fname)]
        row = []
        for j in range(num_frames):
            start = ts + j * (sweep_time + periodicity)
            if not (t_lo <= start <= t_hi):
                continue  # interp1d raises outside the span
            loc = np.array([np.interp(start, loc_times, locs[:, d]) for d in range(locs.shape[1])])
            if j != 0:
                if not row:
                    continue  # the reference's min over an empty row raises
                if np.min(np.linalg.norm(np.array(row)[:, :3] - loc[:3], axis=1)) < thr:
                    continue
            lo = j * samples_per_chirp * num_chirps
            new = data[lo:lo + samples_per_chirp, :]
            if new.shape != (samples_per_chirp, num_lanes):
                continue
            if is_optitrack:
                loc = np.array([-loc[0], loc[2], loc[1]])  # image_processor.py:64-72
            all_locs.append(loc)
            row.append(loc)
            all_data.append(new)
    return {"radar_data": all_data, "poses": np.array(all_locs)}
