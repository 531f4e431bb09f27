"""Raw captures to surface normals, the reference's stage chain
(mmwave_normal_estimation.py:62-68) through this package: .bin captures,
trajectory .npz and experiment JSON on disk -> formats.load_radar_files /
load_robot_loc_files / load_exp_json -> ImageProcessor.correlate_locs_and_meas
-> ImageProcessor.generate_mmwave_surface_normals.  The captures are a
point scatterer's chirps (int16 I/Q, four RX lanes, the TI range offset) seen
from a raster the robot sweeps row by row; checks: every frame's pose, the
caller against the composed NumPy drop-ins, the focus and the normal."""
import json
import pickle

import numpy as np
import pytest

from mmnorm_amd import formats, image_processor as ip, py_image_gpu, py_sn_est_gpu

pytestmark = pytest.mark.gpu

N_ROWS, N_FRAMES, PITCH_X, PITCH_Y = 30, 50, 2e-3, 2.5e-3
PERIOD = 0.003
SPEED = PITCH_X / PERIOD
GRID = dict(min_x=0.160, max_x=0.1805, min_y=0.240, max_y=0.2605, min_z=-0.405,
            max_z=-0.3945, step=0.001)


def build_dataset(root):
    params = ip.params_with(num_frames=N_FRAMES)
    p = ip.get_radar_parameters("77_ghz", False, "almost-square", params)
    lam = ip.C0 / p["min_f"]
    rxo = ip.rx_offsets_for("almost-square", lam)
    xl, yl, zl = ip.imaging_grid(GRID)
    target = np.array([np.float32(xl[10]), np.float32(yl[10]), np.float32(zl[5])], np.float64)
    loc0 = np.array(ip.LOCATION0)
    times, tx = [], []
    names, created, truth = [], [], []
    k = np.arange(p["num_samples"])
    for r in range(N_ROWS):
        t_r = 10.0 + r * 1.0
        y = 0.25 - (N_ROWS - 1) / 2 * PITCH_Y + r * PITCH_Y
        x0 = 0.12
        tt = t_r + np.arange(0, 0.1505, 0.001)
        times.append(tt)
        tx.append(np.stack([x0 + SPEED * (tt - t_r), np.full_like(tt, y), np.zeros_like(tt)], 1)
                  + loc0)
        ts = t_r + 0.0005
        frames = []
        for j in range(N_FRAMES):
            pos = np.array([x0 + SPEED * (ts + j * PERIOD - t_r), y, 0.0])
            truth.append(pos)
            lanes = []
            for o in rxo:
                d = np.linalg.norm(target - pos) + np.linalg.norm(target - pos - o) + 0.15
                lanes.append(3000 * np.exp(2j * np.pi * (p["min_f"] + k * ip.SLOPE_REAL) * d
                                           / ip.C0))
            frames.append(np.stack(lanes, 1))
        fid = 1700000000 + r
        formats.write_adc_bin(f"{root}/adc_data{fid}.bin", np.concatenate(frames, 0))
        names.append(str(fid))
        created.append(ts)
    np.savez(f"{root}/antenna_0.npz", times_77=np.concatenate(times),
             tx_77_locs=np.concatenate(tx), times_24=np.zeros(1), tx_24_locs=np.zeros((1, 3)))
    with open(f"{root}/experiment_info.json", "w") as f:
        json.dump({"aperture size": "almost-square", "rx offset": "x", "background": "table"}, f)
    robot = {"all_radar_filenames": names, "all_radar_created_ts": created}
    # the robot runs as the rig writes them: pickled dictionaries, one per run
    half = len(names) // 2
    for k, (lo, hi) in enumerate(((0, half), (half, len(names)))):
        with open(f"{root}/continuous_{k}.pkl", "wb") as f:
            pickle.dump({"all_radar_filenames": names[lo:hi],
                         "all_radar_created_ts": created[lo:hi], "pattern": 1.0}, f)
    return params, robot, np.array(truth), target, (xl, yl, zl)


def test_captures_to_normals(tmp_path):
    root = str(tmp_path)
    params, robot_written, truth, target, (xl, yl, zl) = build_dataset(root)
    robot = formats.load_robot_files(root)
    assert robot == dict(robot_written, pattern=0.5)
    radar_data = formats.load_radar_files(root)
    robot_loc = formats.load_robot_loc_files(root)
    exp = formats.load_exp_json(root)
    proc = ip.ImageProcessor(False, params=params, grid=GRID)
    data = proc.correlate_locs_and_meas("77_ghz", robot, radar_data, exp, robot_loc,
                                        speed="speed_8")
    assert len(data["poses"]) == N_ROWS * N_FRAMES
    np.testing.assert_allclose(data["poses"][:, :3] - np.array(ip.LOCATION0), truth, atol=1e-9)
    sum_image, axes, normals, centers, meas = proc.generate_mmwave_surface_normals(
        "77_ghz", data, exp)
    assert sum_image.shape == (len(xl), len(yl), len(zl)) and normals.shape[-1] == 3
    # the same stages through the NumPy drop-ins, as image_processor.py:382-396 composes them
    p = ip.get_radar_parameters("77_ghz", False, "almost-square", params)
    lam = ip.C0 / p["min_f"]
    locs = np.asarray(data["poses"])[:, :3] - np.array(ip.LOCATION0)
    clip = np.array(data["radar_data"])[:, :p["num_samples"], :4]
    pad = np.zeros((len(clip), 4096, 4), np.complex64)
    pad[:, :p["num_samples"]] = clip
    md = np.fft.fft(pad, axis=1)
    rxo = ip.rx_offsets_for("almost-square", lam)
    ref_img, ref_c = py_image_gpu.image_cuda(xl, yl, zl, locs, md, rxo, ip.SLOPE_REAL, lam,
                                             p["bandwidth"], p["num_samples"], use_4_rx=True,
                                             is_ti_radar=True, use_interpolated_processing=True)
    ref_n = py_sn_est_gpu.normal_estimates_cuda(ref_img, xl, yl, zl, locs, md, rxo,
                                                ip.SLOPE_REAL, lam, p["bandwidth"],
                                                p["num_samples"], use_4_rx=True,
                                                is_ti_radar=True)
    ref_n = ref_n / np.linalg.norm(ref_n, axis=-1, keepdims=True)
    assert np.max(np.abs(sum_image - ref_img)) <= 1e-4 * np.max(np.abs(ref_img))
    assert np.array_equal(np.isnan(normals), np.isnan(ref_n))
    ok = ~np.isnan(normals[..., 0])
    assert np.degrees(np.arccos(np.clip(np.sum(normals[ok] * ref_n[ok], -1).min(), -1, 1))) < 0.01
    assert np.allclose(centers, ref_c) and np.array_equal(meas, clip)
    # the scatterer's plane peaks at the scatterer, and its vote points up at the aperture
    plane = np.abs(sum_image[:, :, 5])
    assert np.unravel_index(np.argmax(plane), plane.shape) == (10, 10)
    n = normals[10, 10, 5]
    assert not np.isnan(n[0]) and np.degrees(np.arccos(n[2])) < 5.0


def test_normal_estimation_stage_from_the_folder_scheme(tmp_path):
    """mmwave_normal_estimation.py:15-72 as a stage: the same captures laid out
    as the dataset stores them (radar_data/ for the .bin files, the
    experiment JSON in the exp2 folder, an angle folder found by its exp2),
    normal_data.npz out with the reference's keys, equal to the caller run on
    the same files; a second run is skipped."""
    import os
    import shutil
    from mmnorm_amd import normal_estimation
    root = str(tmp_path)
    at = dict(is_los=True, exp_num="2", angles=(0, 90, 0))
    raw = formats.dataset_path(root, "048", "hammer", "77_ghz", **at)
    os.makedirs(raw)
    params, _, _, _, _ = build_dataset(raw)
    os.makedirs(f"{raw}/radar_data")
    for f in os.listdir(raw):
        if f.endswith(".bin"):
            shutil.move(f"{raw}/{f}", f"{raw}/radar_data/{f}")
    shutil.move(f"{raw}/experiment_info.json",
                formats.dataset_path(root, "048", "hammer", "77_ghz", is_json=True, **at)
                + "/experiment_info.json")
    out = normal_estimation.estimate_normals("048", "hammer", True, "_s", False, root=root,
                                             params=params, grid=GRID)
    assert out == formats.dataset_path(root, "048", "hammer", "77_ghz", is_processed=True,
                                       **at) + "/normals/normal_data_s.npz"
    sum_img, axes, ant, ch, normals = formats.load_normal_data(out)
    # the caller on the same files
    proc = ip.ImageProcessor(False, params=params, grid=GRID)
    exp = formats.load_exp_json(formats.dataset_path(root, "048", "hammer", "77_ghz",
                                                     is_json=True, **at))
    data = proc.correlate_locs_and_meas("77_ghz", formats.load_robot_files(raw),
                                        formats.load_radar_files(f"{raw}/radar_data"), exp,
                                        formats.load_robot_loc_files(raw), speed="speed_8")
    want = proc.generate_mmwave_surface_normals("77_ghz", data, exp)
    assert np.array_equal(sum_img, want[0]) and np.array_equal(normals, want[2], equal_nan=True)
    assert all(np.array_equal(a, b) for a, b in zip(axes, want[1]))
    assert np.array_equal(ant, want[3]) and np.array_equal(ch, want[4])
    assert normal_estimation.estimate_normals("048", "hammer", True, "_s", False, root=root,
                                              params=params, grid=GRID) is None
