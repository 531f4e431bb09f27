"""The caller of the hot path, ``ImageProcessor`` of image_processor.py:16-398,
with the reference's class and method signatures so that
mmwave_normal_estimation.py:43,65,68 drops in unchanged:

    processor = ImageProcessor(False)
    data = processor.correlate_locs_and_meas(radar_type, robot_data, radar_data,
                                             exp_data, robot_loc_data, speed="speed_8")
    sum_image, locs, normals, robot_locs, meas = \\
        processor.generate_mmwave_surface_normals(radar_type, data, exp_data)

Radar constants come from ``get_radar_parameters`` (utils/utilities.py:11-26)
over the values of the reference's params.json (``PARAMS``; pass ``params=``
to use another table).  ``generate_mmwave_surface_normals`` keeps every
intermediate on the GPU (the reference passes NumPy arrays between the FFT,
the image and the vote) and returns NumPy arrays, as the reference does.

``interpolate_measurements`` is the reference's zero-padded FFT along the
samples, done by the device FFT on complex64 (the reference: NumPy,
complex128, then narrowed to float32 by the image kernel).
"""
from __future__ import annotations

import copy

import numpy as np
import torch
from scipy import interpolate

from . import _lib, ops, py_image_gpu
from .py_image_gpu import MAX_GROUP_ROWS

C0 = 2.99792458e8                        # scipy.constants.c, image_processor.py:5
LOCATION0 = (0.42, -0.23, 0.23)          # image_processor.py:317
OPTITRACK_ORIGIN = (-1.33640909, -0.87117398, 1.17517328)   # image_processor.py:341
GRID = dict(min_x=-0.1, max_x=0.45, min_y=-0.05, max_y=0.55, min_z=-0.5, max_z=-0.25,
            step=0.001)                  # image_processor.py:327-336
SLOPE_REAL = 58.6116 * 1e6 * 1e6 / 10e6  # image_processor.py:358
TX_OFFSET = 0.005                        # image_processor.py:359

# The values of the reference's utils/params.json (the processing flags and
# the 77 GHz rig's radar constants).
PARAMS = {
    "processing": {"use_cuda": True, "use_interpolated_processing": True,
                   "path_to_data": "sample_data"},
    "robot_collected": {"77_ghz": {"almost-square": {
        "min_f": 77.5001e9, "max_f": 80.501e9, "bandwidth": 3.0009e9, "slope": 5861160,
        "num_samples": 512, "num_rx_antenna": 4, "num_chirps": 1, "periodicity": 0.003,
        "num_frames": 4000, "sweep_time": 0}}},
}


def get_radar_parameters(radar_type="77_ghz", is_sim=False, aperture_type="normal",
                         params=None):
    """The radar constants of one rig (utils/utilities.py:11-26): 'large'
    apertures use the 'normal' entry; real-data entries are per aperture type.
    A missing entry raises KeyError, as the reference's dictionary lookups do."""
    if aperture_type == "large":
        aperture_type = "normal"
    p = PARAMS if params is None else params
    current = p["simulation" if is_sim else "robot_collected"][radar_type]
    if not is_sim:
        current = current[aperture_type]
    return current


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
    """The steps of image_processor.py:296-398 from arrays: the SAR image, the
    grid axes, unit surface normals (NaN outside the 5 % filter), the rows'
    phase centres and the clipped measurements, from the correlated poses
    (N, >= 3; their first three columns minus location0 are the radar-frame
    positions) and measurements (N, >= num_samples, >= R).  With is_sim the
    slope defaults to bandwidth / (num_samples - 1) (the reference's branch
    reads an undefined name there, image_processor.py:375)."""
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


class ImageProcessor:
    """image_processor.py:16-398 on the GPU.

    is_optitrack, radar_type: as the reference's constructor.  Keyword-only
    additions: params (the parameter table, default PARAMS; its 'processing'
    entry supplies use_interpolated_processing as the reference's params.json
    does), use_interpolated_processing (overrides that flag), grid (the
    imaging grid of generate_mmwave_surface_normals, default the reference's
    1 mm box) and K (re-anchor interval of the full-frequency recurrence).
    """

    def __init__(self, is_optitrack=False, radar_type="77_ghz", *, params=None,
                 use_interpolated_processing=None, grid=None, K=64):
        self.is_optitrack = is_optitrack
        self.radar_type = radar_type
        self.params = PARAMS if params is None else params
        if use_interpolated_processing is None:
            use_interpolated_processing = bool(
                self.params.get("processing", {}).get("use_interpolated_processing", True))
        self.use_interpolated_processing = use_interpolated_processing
        self.grid = GRID if grid is None else grid
        self.K = K

    # ------------------------------------------------------------ helpers --
    def interpolate_measurements(self, radar_data, FFT_SIZE=4096):
        """image_processor.py:22-35 (NumPy in and out, the FFT on the GPU)."""
        if np.asarray(radar_data).shape[1] > FFT_SIZE:
            raise AssertionError("more samples than FFT_SIZE")
        return interpolate_measurements(radar_data, FFT_SIZE)

    def get_interpolated_funct(self, transform_times, transforms):
        """One linear interpolant per column (image_processor.py:37-49)."""
        transforms = np.asarray(transforms)
        return [interpolate.interp1d(transform_times, transforms[:, i])
                for i in range(transforms.shape[1])]

    def get_interpolated_transforms_from_funct(self, functs, dist_times):
        """image_processor.py:51-62."""
        out = np.zeros((len(dist_times), len(functs)))
        for i, f in enumerate(functs):
            out[:, i] = f(dist_times)
        return out

    def convert_opt_frame(self, loc):
        """image_processor.py:64-72."""
        return np.array([-loc[0], loc[2], loc[1]])

    # ------------------------------------------------------- correlation --
    def correlate_locs_and_meas(self, radar_type, robot_data, radar_data, exp_data,
                                robot_loc_data=None, speed="speed_8"):
        """The pose of every radar frame (image_processor.py:74-154): frame j of
        radar file i starts at created_ts + j (sweep_time + periodicity) and
        takes the trajectory's linear interpolation there; frames outside the
        trajectory's time span are dropped (the interpolant raises), so are
        frames within 3/16 of a wavelength of a frame kept earlier in the same
        file, and frames with fewer samples than a chirp.  As in the reference,
        once a file's first frame is dropped its later frames are dropped too
        (the distance test over an empty row raises).  ``speed`` is unused, as
        there.  Returns {'radar_data': list of (samples, lanes) arrays,
        'poses': (n, D)}."""
        if radar_type not in ("77_ghz", "24_ghz"):
            raise AssertionError("Please choose a valid radar type.")
        aperture_type = exp_data["aperture size"]
        p = get_radar_parameters(radar_type=radar_type, is_sim=False,
                                 aperture_type=aperture_type, params=self.params)
        n_frames, sweep, period = p["num_frames"], p["sweep_time"], p["periodicity"]
        spc, n_chirp = p["num_samples"], p["num_chirps"]
        thr = (C0 / p["min_f"]) * (3 / 16)
        lanes = 2 if radar_type == "24_ghz" else 4
        tag = radar_type[:2]
        times = np.asarray(robot_loc_data[f"times_{tag}"])
        functs = self.get_interpolated_funct(times, robot_loc_data[f"tx_{tag}_locs"])
        t_lo, t_hi = np.min(times), np.max(times)
        all_locs, all_data = [], []
        for i, filename in enumerate(robot_data["all_radar_filenames"]):
            created_ts = robot_data["all_radar_created_ts"][i]
            if created_ts == -1 or int(filename) not in radar_data:
                continue
            data = radar_data[int(filename)]
            starts = created_ts + np.arange(n_frames) * (sweep + period)
            inside = (starts >= t_lo) & (starts <= t_hi)
            locs = np.full((n_frames, len(functs)), np.nan)
            if inside.any():   # the interpolants are elementwise: one call per column
                locs[inside] = self.get_interpolated_transforms_from_funct(functs,
                                                                           starts[inside])
            row = []
            for j in range(n_frames):
                if not inside[j]:
                    continue
                loc = locs[j]
                if j != 0:
                    if not row:
                        continue
                    if np.min(np.linalg.norm(np.array(row)[:, :3] - loc[:3], axis=1)) < thr:
                        continue
                lo = j * spc * n_chirp
                new = data[lo:lo + spc, :]
                if new.shape != (spc, lanes):
                    continue
                if self.is_optitrack:
                    loc = self.convert_opt_frame(loc)
                all_locs.append(loc)
                row.append(loc)
                all_data.append(new)
        return {"radar_data": all_data, "poses": np.array(all_locs)}

    # -------------------------------------------------------------- images --
    def generate_sar_image_radar_frame(self, radar_type, data, exp_data, data_bg=None,
                                       plot=False, bg_chirp_num=0):
        """The SAR image without normals (image_processor.py:156-294): optional
        background subtraction, radar-frame positions, the grid of the rig and
        the background type, RX offsets along exp_data['rx offset'], the image
        divided by the number of poses.  Returns (image, (x, y, z),
        radar-frame positions).  Differences, by intent: data['poses'] is not
        shifted in place (the reference subtracts location0 from the caller's
        array); ``plot`` is ignored; the reference's GPU branch divides the
        (image, centres) tuple image_cuda returns, which fails, where this
        divides the image."""
        aperture_type = exp_data["aperture size"]
        rx_dir = exp_data["rx offset"]
        is_tripod = exp_data["background"] == "tripod"
        locs = np.array(data["poses"], dtype=np.float64, copy=True)
        radar_data = np.array(data["radar_data"])
        p = get_radar_parameters(radar_type=radar_type, is_sim=False,
                                 aperture_type=aperture_type, params=self.params)
        S = p["num_samples"]
        wavelength = C0 / p["min_f"]
        slope, bandwidth = p["slope"], p["bandwidth"]
        if data_bg is not None:
            bg = np.array(data_bg["radar_data"])[bg_chirp_num, :S, :]
            radar_data = radar_data.copy()
            radar_data[:, :S, :] -= bg
        origin = np.array(OPTITRACK_ORIGIN if self.is_optitrack else LOCATION0)
        locs[:, :3] -= origin
        if radar_type == "77_ghz":
            bx, by = (-0.1, 0.45), (-0.05, 0.55)
        else:
            bx, by = (0 - 0.05, np.max(locs[:, 0]) + 0.05), (0.05, 0.65)
        bz = (-0.37, -0.15) if is_tripod else (-0.55, -0.3)
        xl, yl, zl = (np.arange(lo, hi, 0.001) for lo, hi in (bx, by, bz))
        if radar_type == "24_ghz":
            rxo = np.array([[0, -0.021, 0], [0, -0.021 - 3 / 240 / 2, 0]])
            max_rx = 2
        else:
            r = wavelength / 2
            steps = [TX_OFFSET + r * 3, TX_OFFSET + r * 2, TX_OFFSET + r, TX_OFFSET]
            if rx_dir == "x":
                rxo = np.array([[-s, 0, 0] for s in steps])
            elif rx_dir == "y":
                rxo = np.array([[0, -s, 0] for s in steps])
            else:
                raise ValueError("rx_offset_dir must be x or y")
            max_rx = 4
        clip = radar_data[:, :S, :max_rx]
        md = self.interpolate_measurements(clip) if self.use_interpolated_processing else clip
        image = py_image_gpu.image_cuda(
            xl, yl, zl, locs[:, :3], md, rxo, slope, wavelength, bandwidth, S,
            use_4_rx=max_rx != 1, is_ti_radar=radar_type == "77_ghz",
            use_interpolated_processing=self.use_interpolated_processing,
            return_phase_centers=False, K=self.K)
        image = image / len(locs)
        return image, (xl, yl, zl), locs

    def generate_mmwave_surface_normals(self, radar_type, data, exp_data, is_sim=False):
        """image_processor.py:296-398: returns (sum_image, (x, y, z), unit
        normals (NaN outside the 5 % filter), phase centres, clipped
        measurements).  data = {'poses', 'radar_data'} as
        correlate_locs_and_meas gives; exp_data['aperture size'] picks the
        radar constants and the RX layout."""
        aperture_type = exp_data["aperture size"]
        p = get_radar_parameters(radar_type=radar_type, is_sim=False,
                                 aperture_type=aperture_type, params=self.params)
        poses = np.asarray(data["poses"], dtype=np.float64)
        origin = OPTITRACK_ORIGIN if self.is_optitrack else LOCATION0
        slope = (p["max_f"] - p["min_f"]) / (p["num_samples"] - 1) if is_sim else None
        return generate_mmwave_surface_normals(
            poses, np.array(data["radar_data"]), aperture_type, p["min_f"], p["num_samples"],
            p["bandwidth"], slope=slope, is_sim=is_sim,
            use_interpolated_processing=self.use_interpolated_processing, location0=origin,
            grid=self.grid, K=self.K)


def params_with(**overrides):
    """A deep copy of PARAMS with the 77 GHz 'almost-square' constants
    overridden (small captures for synthetic runs)."""
    p = copy.deepcopy(PARAMS)
    p["robot_collected"]["77_ghz"]["almost-square"].update(overrides)
    return p
