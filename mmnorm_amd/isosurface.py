"""Drop-ins for the RSDF consumer: ``ray_tracing.ray_matching``
(ray_tracing.cpp:34-76) and the isosurface search of
``isosurface_optimization.py`` (channel simulation :111-209, loss :211-246,
channel filter :248-290, tau selection :292-358, point cloud :360-377, the
search loop :475-521).  Same names, arguments and results.

The per-ray work -- ray generation, sphere tracing, the refine step, the
round-trip distance and the sum of exp(+j 2 pi d / lambda) over hits -- runs
in one pass of libmmnorm_hip.so per antenna and component
(``ops.simulate_channel``), and the range-profile loss of a whole batch of
antennas in one launch (``ops.channel_loss``; an exact-twiddle DFT of the
selected bins instead of the reference's full N x N DFT).  What stays on
the host is what the reference does once per search with NumPy/SciPy: the
channel filter, the peak picking of the tau selection and the point cloud.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import _lib, ops

C0 = 2.99792458e8
UNDEF_VAL = -50  # isosurface_optimization.py:26

# Range window of the loss (isosurface_optimization.py:437-448); the
# reference sets these globals in its main block, here set_range_window does.
NUM_SAMPLES = None
FFT_SIZE = 512
DISTANCES = None
MAX_DIST_IDX = None
MAX_DIST_IDX_REAL = None
MIN_DIST_IDX = None


def set_range_window(num_samples, wavelengths, fft_size=512, min_dist=0.35,
                     max_dist_all=0.55, max_dist_real=0.55):
    """The loss's FFT size and range bins from the chirp (isosurface_optimization.py:437-448)."""
    global NUM_SAMPLES, FFT_SIZE, DISTANCES, MAX_DIST_IDX, MAX_DIST_IDX_REAL, MIN_DIST_IDX
    wl = np.asarray(wavelengths, dtype=np.float64)[:num_samples]
    NUM_SAMPLES = int(num_samples)
    FFT_SIZE = int(fft_size)
    f_max = 3e8 / wl[-1]
    DISTANCES = np.arange(0, FFT_SIZE) * 3e8 / (2 * (f_max - 77.5e9)) * NUM_SAMPLES / FFT_SIZE
    MAX_DIST_IDX = int(np.count_nonzero(DISTANCES < max_dist_all))
    MAX_DIST_IDX_REAL = int(np.count_nonzero(DISTANCES < max_dist_real))
    MIN_DIST_IDX = int(np.count_nonzero(DISTANCES < min_dist))


def _freqs(wavelengths, n):
    """(f0, df) with c / wavelengths[j] = f0 + j df, as the reference's
    wavelengths are built (sdf.py:249)."""
    wl = np.asarray(wavelengths.detach().cpu() if torch.is_tensor(wavelengths) else wavelengths,
                    dtype=np.float64)[:n]
    f = C0 / wl
    if n == 1:
        return float(f[0]), 0.0
    f0, df = float(f[0]), float((f[-1] - f[0]) / (n - 1))
    if np.max(np.abs(f - (f0 + df * np.arange(n)))) > 1e-6 * abs(f0):
        raise ValueError("wavelengths must be c / (f0 + j df) (a linear chirp)")
    return f0, df


def _scalar(x):
    return float(x.item()) if torch.is_tensor(x) else float(x)


def ray_matching(w_h_3, w_h, grid, width, height, width_fov, height_fov,
                 bounding_box_min_x, bounding_box_min_y, bounding_box_min_z,
                 bounding_box_max_x, bounding_box_max_y, bounding_box_max_z,
                 grid_res_x, grid_res_y, grid_res_z, eye_x, eye_y, eye_z,
                 rx_dir_x, rx_dir_y, rx_dir_z):
    """Drop-in for ray_tracing.ray_matching: [intersection_pos, voxel_position,
    directions], each float32 (width, height, 3); w_h_3 and w_h only give sizes,
    as in the reference."""
    g = grid if grid.dtype == torch.float32 else grid.float()
    g = g.contiguous()
    if tuple(g.shape) != (int(grid_res_x), int(grid_res_y), int(grid_res_z)):
        raise ValueError("grid shape and grid_res disagree")
    bb = [bounding_box_min_x, bounding_box_min_y, bounding_box_min_z,
          bounding_box_max_x, bounding_box_max_y, bounding_box_max_z]
    hit, vox, dirs = ops.ray_matching(g, int(width), int(height), _scalar(width_fov),
                                      _scalar(height_fov), [_scalar(v) for v in bb],
                                      [_scalar(eye_x), _scalar(eye_y), _scalar(eye_z)],
                                      [_scalar(rx_dir_x), _scalar(rx_dir_y), _scalar(rx_dir_z)])
    return [hit, vox, dirs]


def _simulate_channels(bounding_box, bounding_box_lower, voxel_size, grid_res_x, grid_res_y,
                       grid_res_z, width, height, width_fov, height_fov, grid, rx_loc, rx_dir,
                       wavelengths, taus, tau_bounds, is_real_data, apply_tau_conversion=True,
                       sdf_indices=None):
    """Channel (complex64, one per wavelength) of antenna ``rx_loc`` from the
    isosurfaces grid[..., k] - taus[k] of the selected components
    (isosurface_optimization.py:111-209).  ``bounding_box_lower`` is implied
    by ``bounding_box`` and not needed."""
    dev = grid.device
    num_sdfs = grid.shape[-1]
    n_wl = int(wavelengths.shape[0])
    f0, df = _freqs(wavelengths, n_wl)
    rx = [float(v) for v in (rx_loc.detach().cpu().numpy() if torch.is_tensor(rx_loc)
                             else np.asarray(rx_loc))]
    eye = [float(np.float32(v)) for v in rx]
    rd = [_scalar(v) for v in rx_dir]
    bb = [float(v) for v in bounding_box]
    vsz = float(np.float32(voxel_size.item() if torch.is_tensor(voxel_size) else voxel_size))
    combined = torch.zeros(n_wl, dtype=torch.complex64, device=dev)
    if sdf_indices is None:
        sdf_indices = range(num_sdfs)
    for k in sdf_indices:
        k = int(k)
        tau = taus[k]
        if apply_tau_conversion:
            lo, hi = tau_bounds[k]
            tau = torch.clip(torch.as_tensor(tau, dtype=torch.float32),
                             torch.as_tensor(lo, dtype=torch.float32),
                             torch.as_tensor(hi, dtype=torch.float32))
        sub = grid[..., k]
        sub = sub.float().contiguous() if (sub.dtype != torch.float32 or
                                           not sub.is_contiguous()) else sub
        combined += ops.simulate_channel(sub, float(np.float32(_scalar(tau))), int(width),
                                         int(height), float(width_fov), float(height_fov), bb,
                                         eye, rd, rx, vsz, f0, df, n_wl, is_ti=is_real_data)
    return combined


def _host_rows(a):
    return np.asarray(a.detach().cpu().numpy() if torch.is_tensor(a) else a,
                      dtype=np.float64).reshape(-1, 3)


def _channel_prefix(bounding_box, width, height, width_fov, height_fov, grid_cm, rx_locs,
                    rx_dirs, tau_bounds):
    """ops.ray_prefix for the simulations of _simulate_channel_batch (antenna
    r, component k in that order) and every offset the clip to ``tau_bounds``
    can give (>= the smallest lower bound)."""
    n_comp = grid_cm.shape[0]
    rx, rd = _host_rows(rx_locs), _host_rows(rx_dirs)
    tau_lo = min(float(np.float32(_scalar(torch.as_tensor(tau_bounds[k][0],
                                                          dtype=torch.float32))))
                 for k in range(n_comp))
    return ops.ray_prefix(grid_cm, np.tile(np.arange(n_comp, dtype=np.int32), len(rx)),
                          np.repeat(rx.astype(np.float32), n_comp, axis=0),
                          np.repeat(rd.astype(np.float32), n_comp, axis=0), int(width),
                          int(height), float(width_fov), float(height_fov),
                          [float(v) for v in bounding_box], tau_lo)


def _simulate_channel_batch(bounding_box, voxel_size, width, height, width_fov, height_fov,
                            grid_cm, rx_locs, rx_dirs, wavelengths, taus, tau_bounds,
                            is_real_data, prefix=None):
    """_simulate_channels for several antennas at once: row r is the combined
    channel of antenna (rx_locs[r], rx_dirs[r]) over every component of
    ``grid_cm`` (n_comp, NX, NY, NZ) with the clipped offsets ``taus``, bit for
    bit what _simulate_channels returns for it (the same per-simulation values,
    the components added to zeros in order), from one batched launch
    (ops.simulate_channels), or, given the ``prefix`` of _channel_prefix for
    the same antennas, from its live rays (ops.simulate_channels_prefixed)."""
    n_comp = grid_cm.shape[0]
    n_wl = int(wavelengths.shape[0])
    f0, df = _freqs(wavelengths, n_wl)
    rx, rd = _host_rows(rx_locs), _host_rows(rx_dirs)
    n_ant = len(rx)
    tv = np.empty(n_comp, dtype=np.float32)
    for k in range(n_comp):
        lo, hi = tau_bounds[k]
        tv[k] = float(np.float32(_scalar(torch.clip(
            torch.as_tensor(taus[k], dtype=torch.float32),
            torch.as_tensor(lo, dtype=torch.float32), torch.as_tensor(hi, dtype=torch.float32)))))
    comp = np.tile(np.arange(n_comp, dtype=np.int32), n_ant)
    rep = lambda a: np.repeat(a, n_comp, axis=0)
    vsz = float(np.float32(voxel_size.item() if torch.is_tensor(voxel_size) else voxel_size))
    if prefix is not None:
        sims = ops.simulate_channels_prefixed(prefix, np.tile(tv, n_ant), rep(rx), vsz, f0, df,
                                              n_wl, is_ti=is_real_data)
    else:
        sims = ops.simulate_channels(grid_cm, comp, np.tile(tv, n_ant),
                                     rep(rx.astype(np.float32)), rep(rd.astype(np.float32)),
                                     rep(rx), int(width), int(height), float(width_fov),
                                     float(height_fov), [float(v) for v in bounding_box], vsz,
                                     f0, df, n_wl, is_ti=is_real_data)
    sims = sims.view(n_ant, n_comp, n_wl)
    combined = torch.zeros((n_ant, n_wl), dtype=torch.complex64, device=grid_cm.device)
    for k in range(n_comp):
        combined += sims[:, k]
    return combined


def loss_fn(output_ch, target_chs, is_real_data):
    """Range-profile loss of one simulated / measured channel pair
    (isosurface_optimization.py:211-246); needs set_range_window first."""
    return batch_loss(output_ch[None], target_chs[None], is_real_data)[0]


def batch_loss(output_chs, target_chs, is_real_data):
    """loss_fn for B pairs at once: complex (B, F) each -> float32 (B,)."""
    if MIN_DIST_IDX is None:
        raise RuntimeError("call set_range_window(num_samples, wavelengths) first")
    hi_t = MAX_DIST_IDX_REAL if is_real_data else FFT_SIZE
    return ops.channel_loss(output_chs.to(torch.complex64), target_chs.to(torch.complex64),
                            FFT_SIZE, MIN_DIST_IDX, MAX_DIST_IDX, min(hi_t, MAX_DIST_IDX))


def _filter_channels(channels, rx_locs, rx_dirs, is_real_data=True):
    """Keep antennas whose range profile has a strong, interior peak
    (isosurface_optimization.py:248-290), relaxing the thresholds twice if
    fewer than 20 remain."""
    ch = np.asarray(channels.detach().cpu())
    # the reference's FFT length is its sample count, which equals FFT_SIZE
    # there (512); n = FFT_SIZE keeps the bins right for other sample counts
    prof = np.abs(np.fft.fft(ch, n=FFT_SIZE, axis=1))
    if is_real_data:
        prof[:, MAX_DIST_IDX_REAL:] = 0
    prof = prof[:, MIN_DIST_IDX:MAX_DIST_IDX]
    peak = np.max(prof, axis=1)
    arg = np.argmax(prof, axis=1)
    interior = np.logical_and(arg != 0, arg != prof.shape[1] - 1)
    peaky = np.sum(prof, axis=1) / prof.shape[1] < peak * 0.7
    keep = None
    for frac in (0.5, 0.1, 0.05):
        keep = np.logical_and(np.logical_and(peak > frac * np.max(peak), peaky), interior)
        if np.count_nonzero(keep) >= 20:
            break
    sel = torch.from_numpy(keep).to(channels.device)
    return channels[sel, :], rx_locs[sel.to(rx_locs.device), :], rx_dirs[sel.to(rx_dirs.device), :]


def _select_tau(all_taus, all_losses):
    """The isosurface offsets: the loss minimum, or a stronger-looking peak of
    the flipped loss closer to the top surface (isosurface_optimization.py:292-358)."""
    from scipy.signal import find_peaks
    if all_losses.ndim == 1:
        imin = int(np.argmin(all_losses))
        thr = 0.3 if imin in (0, len(all_losses) - 1) else 0.9
        flipped = np.nanmax(all_losses) - all_losses
        peaks, _ = find_peaks(flipped, height=flipped[imin] * thr, prominence=1)
        peaks = list(peaks) if len(peaks) else [imin]
        if imin not in peaks and imin != 0 and imin != len(all_losses) - 1:
            peaks = sorted(peaks + [imin])
        return all_taus[peaks[-1]]
    best = list(np.unravel_index(np.argmin(all_losses), all_losses.shape))
    for i in range(len(best)):
        line = all_losses
        for j in range(len(best)):
            if j < i:
                line = line[best[j]]
            elif j > i:
                line = line[:, best[j]]
        line = np.nanmax(line) - line
        height = max(max(line) * 0.4, 2)
        peaks, _ = find_peaks(line, height=height, prominence=0.1)
        best[i] = peaks[-1] if len(peaks) else best[i]
    taus = all_taus
    for i in best:
        taus = taus[i]
    return taus


def _get_final_point_cloud(sdf, taus, coords):
    """Voxel coordinates within 1 mm of each chosen isosurface
    (isosurface_optimization.py:360-377)."""
    pts = [coords[np.abs(sdf[..., k] - taus[k]) < 0.001, :] for k in range(sdf.shape[-1])]
    return np.concatenate(pts, axis=0) if pts else np.zeros((0, 3))


def prepare_components(rsdf, normals, component_labels, valid_labels):
    """Per-component SDF grids with UNDEF_VAL outside each component, and the
    tau bounds (min + 15 mm, max - 15 mm) (isosurface_optimization.py:75-81).
    Returns (sdf float32 (X,Y,Z,n), tau_bounds float32 (n, 2)) as device tensors."""
    dev = torch.device("cuda", torch.cuda.current_device())
    flat = torch.as_tensor(np.asarray(rsdf), dtype=torch.float32, device=dev).clone()
    labels = torch.as_tensor(np.asarray(component_labels).astype(np.int32), device=dev)
    valid = [int(v) for v in np.asarray(valid_labels) if int(v) != 0]
    flat[torch.isnan(flat)] = UNDEF_VAL
    sdf = flat[..., None].repeat(1, 1, 1, len(valid))
    bounds = torch.zeros(len(valid), 2, dtype=torch.float32)
    for i, lab in enumerate(valid):
        sdf[..., i][labels != lab] = UNDEF_VAL
        sel = flat[torch.logical_and(labels == lab, flat > UNDEF_VAL)]
        bounds[i, 0] = sel.min() + 0.015
        bounds[i, 1] = sel.max() - 0.015
    return sdf.contiguous(), bounds


def search_taus(sdf, rx_locs, rx_dirs, channels, bounding_box, wavelengths, voxel_size,
                tau_bounds, num_samples, is_real_data=True, width=2048, height=2048,
                width_fov=80, height_fov=80, num_rx=50, generator=None, skip=15,
                min_dist=0.35, max_dist=0.55):
    """Grid search of the isosurface offsets (isosurface_optimization.py:454-521):
    taus on a grid (step lambda_0 / 2, or 2 lambda_0 with 3+ components, the
    first ``skip`` values of the first axis dropped), the loss summed over
    ``num_rx`` randomly chosen antennas for each, then _select_tau.  Returns
    (taus, tau_grid, all_losses).  min_dist / max_dist set the range window
    (the reference's 0.35 / 0.55 m suit its rig)."""
    set_range_window(num_samples, np.asarray(wavelengths.detach().cpu()
                                             if torch.is_tensor(wavelengths) else wavelengths),
                     min_dist=min_dist, max_dist_all=max_dist, max_dist_real=max_dist)
    wl = wavelengths[:num_samples]
    channels, rx_locs, rx_dirs = _filter_channels(channels, rx_locs, rx_dirs, is_real_data)
    n_sdf = sdf.shape[-1]
    step = float(wl[0]) / 2 if n_sdf < 3 else float(wl[0]) * 2
    vectors = [torch.arange(float(tau_bounds[i, 0]), float(tau_bounds[i, 1]), step,
                            dtype=torch.float32) for i in range(n_sdf)]
    tau_grid = torch.stack(torch.meshgrid(*vectors, indexing="ij"), dim=-1)[skip:, ...]
    shape = tau_grid.shape[:-1]
    order = torch.randperm(len(rx_locs), generator=generator)[:num_rx]
    dev = sdf.device
    tgt = channels[order].to(dev)
    # every antenna and component of one offset in one batched simulation
    grid_cm = sdf.permute(3, 0, 1, 2).float().contiguous()
    sel = order.tolist()
    locs = (rx_locs.detach().cpu().numpy() if torch.is_tensor(rx_locs)
            else np.asarray(rx_locs))[sel]
    dirs = (rx_dirs.detach().cpu().numpy() if torch.is_tensor(rx_dirs)
            else np.asarray(rx_dirs))[sel]
    # the march up to the first offset-dependent step, once for the search
    prefix = _channel_prefix(bounding_box, width, height, width_fov, height_fov, grid_cm,
                             locs, dirs, tau_bounds)
    # the offsets' losses stay on the device until the loop ends (one read-back)
    n_flat = int(np.prod(shape))
    losses = torch.zeros(n_flat, dtype=torch.float32, device=dev)
    for flat_i in range(n_flat):
        taus = tau_grid[np.unravel_index(flat_i, shape)]
        sims = _simulate_channel_batch(bounding_box, voxel_size, width, height, width_fov,
                                       height_fov, grid_cm, locs, dirs, wl, taus, tau_bounds,
                                       is_real_data, prefix=prefix)
        losses[flat_i] = batch_loss(sims, tgt, is_real_data).sum()
    losses = losses.cpu().reshape(shape)
    taus = _select_tau(tau_grid.numpy(), losses.numpy())
    return taus, tau_grid, losses


def read_sdf_file(path):
    """_read_sdf_file (isosurface_optimization.py:30-83) on the stage file at
    ``path`` (the sdf.npz of sdf.py:226-253): channels scaled by 1e-5/6,
    per-component grids with UNDEF_VAL outside each component and the tau
    bounds (prepare_components), the antennas, look directions (float32),
    channels and wavelengths on the device.  The per-component normals the
    reference also forms are not used by the search and are not built."""
    from . import formats
    data = formats.load_sdf_file(path)
    if data is None:
        raise FileNotFoundError(path)
    dev = torch.device("cuda", torch.cuda.current_device())
    sdf, tau_bounds = prepare_components(data["rsdf"], data["normals"],
                                         data["component_labels"], data["valid_labels"])
    channels = torch.from_numpy(np.asarray(data["channels"]))
    channels = channels * (0.00001 / 6)
    return {
        "sdf": sdf, "tau_bounds": tau_bounds,
        "rx_locs": torch.from_numpy(np.asarray(data["rx_locs"])).to(dev),
        "rx_dirs": torch.from_numpy(np.asarray(data["rx_dirs"])).to(torch.float32).to(dev),
        "channels": channels.to(dev),
        "bounding_box": data["bounding_box"],
        "wavelengths": torch.from_numpy(np.asarray(data["wavelengths"])).to(dev),
        "voxel_size": data["voxel_size"], "num_samples": int(data["num_samples"]),
        "coords": data["coords"], "sum_img": data["sum_img"],
    }


def run_stage(sdf_path, out_dir=None, ext="", is_real_data=True, **search_kw):
    """The isosurface stage of isosurface_optimization.py:379-540 on a stage
    file: read it, search the offsets (search_taus with the reference's
    2048^2 rays, +-80 degrees, 50 antennas unless overridden), keep the voxels
    within 1 mm of the chosen isosurfaces, and write
    opt_output_optimization_result<ext>.npz (taus, points) and
    opt_output_full_data<ext>.npz (sdf, coords, sum_img, num_samples,
    all_taus, all_losses) to ``out_dir``, as the reference's
    save_optimization_file calls do.  Returns the two dictionaries."""
    from . import formats
    d = read_sdf_file(sdf_path)
    taus, tau_grid, losses = search_taus(
        d["sdf"], d["rx_locs"], d["rx_dirs"], d["channels"], d["bounding_box"],
        d["wavelengths"], d["voxel_size"], d["tau_bounds"], d["num_samples"],
        is_real_data=is_real_data, **search_kw)
    sdf_host = d["sdf"].cpu().numpy()
    result = {"taus": taus, "points": _get_final_point_cloud(sdf_host, taus, d["coords"])}
    full = {"sdf": sdf_host, "coords": d["coords"], "sum_img": d["sum_img"],
            "num_samples": d["num_samples"], "all_taus": tau_grid.numpy(),
            "all_losses": losses.numpy()}
    if out_dir is not None:
        formats.save_optimization_file(
            os.path.join(out_dir, f"opt_output_optimization_result{ext}.npz"), result)
        formats.save_optimization_file(
            os.path.join(out_dir, f"opt_output_full_data{ext}.npz"), full)
    return result, full


def main(argv=None):
    """Command line of isosurface_optimization.py:379-395: the stage file and
    the output folder from the dataset's folder scheme (--root, --id, --name,
    --is_los, --ext; experiment 2 as there), or given directly (--sdf,
    --out)."""
    import argparse
    from . import formats
    ap = argparse.ArgumentParser(description="isosurface offset search on the GPU")
    ap.add_argument("--ext", default="tmp")
    ap.add_argument("--id", default="tmp")
    ap.add_argument("--name", default="tmp")
    ap.add_argument("--is_los", default="y")
    ap.add_argument("--overwrite_existing", default="n")
    ap.add_argument("--root", default=".", help="folder holding the dataset")
    ap.add_argument("--sdf", default=None, help="stage file (instead of the folder scheme)")
    ap.add_argument("--out", default=None, help="output folder (with --sdf)")
    ap.add_argument("--rays", type=int, default=2048, help="rays per side per antenna")
    a = ap.parse_args(argv)
    if a.sdf is None:
        angles = formats.find_obj_angles(a.root, a.id, a.name, "2")
        proc = formats.dataset_path(a.root, a.id, a.name, "77_ghz", is_sim=False,
                                    is_los=a.is_los in ("y", "Y"), exp_num="2",
                                    angles=angles, is_processed=True)
        sdf_path, out = f"{proc}/sdf/sdf{a.ext}.npz", f"{proc}/optimization"
    else:
        sdf_path, out = a.sdf, a.out
    done = out and os.path.exists(os.path.join(out, f"opt_output_optimization_result{a.ext}.npz"))
    if done and a.overwrite_existing not in ("y", "Y"):
        print("The output file already exists and overwrite_existing is not set: skipping.")
        return 0
    result, _ = run_stage(sdf_path, out, ext=a.ext, width=a.rays, height=a.rays)
    print(f"taus {np.asarray(result['taus']).tolist()}, {len(result['points'])} points")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
