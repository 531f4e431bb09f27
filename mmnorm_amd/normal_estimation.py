"""The normal-estimation stage of mmwave_normal_estimation.py:15-72 on the
GPU: a real-data experiment's raw captures (experiment 2, the 77 GHz radar)
to processed/.../normals/normal_data<ext>.npz.

    python -m mmnorm_amd.normal_estimation --root <dataset root> --id 048 --name hammer

The loader steps (generic_loader.py:632-650: the experiment JSON, the
robot-run records, the .bin captures under radar_data/, the trajectory
files) come from ``formats``; the pose correlation and the image, vote and
unit normals from ``image_processor.ImageProcessor`` (correlate_locs_and_meas
with speed "speed_8", generate_mmwave_surface_normals); the output has the
keys of generic_loader.save_surface_normals.
"""
from __future__ import annotations

import os
import time

from . import formats
from .image_processor import ImageProcessor


def estimate_normals(obj_id, name, is_los=True, ext="", overwrite_existing=False, root=".",
                     path_to_data="sample_data", params=None, grid=None):
    """Run the stage for one object; returns the output path, or None when
    the output exists and ``overwrite_existing`` is false (the reference's
    skip)."""
    radar_type, exp_num = "77_ghz", "2"
    angles = formats.find_obj_angles(root, obj_id, name, exp_num, path_to_data)
    at = dict(is_sim=False, is_los=is_los, exp_num=exp_num, angles=angles,
              path_to_data=path_to_data)
    out = (formats.dataset_path(root, obj_id, name, radar_type, is_processed=True, **at)
           + f"/normals/normal_data{ext}.npz")
    if not overwrite_existing and os.path.exists(out):
        print("The output file already exists and overwrite_existing is not set: skipping.")
        return None
    t1 = time.time()
    raw = formats.dataset_path(root, obj_id, name, radar_type, **at)
    exp_data = formats.load_exp_json(
        formats.dataset_path(root, obj_id, name, radar_type, is_json=True, **at))
    robot_data = formats.load_robot_files(raw)
    radar_data = formats.load_radar_files(os.path.join(raw, "radar_data"), radar_type)
    robot_loc_data = formats.load_robot_loc_files(raw)
    processor = ImageProcessor(False, params=params, grid=grid)
    data = processor.correlate_locs_and_meas(radar_type, robot_data, radar_data, exp_data,
                                             robot_loc_data, speed="speed_8")
    sum_image, locs, normals, robot_locs, meas = processor.generate_mmwave_surface_normals(
        radar_type, data, exp_data)
    formats.save_normal_data(out, sum_image, *locs, robot_locs, meas, normals)
    print(f"Took {time.time() - t1:.2f} sec to finish processing image")
    return out


def main(argv=None):
    """The command line of mmwave_normal_estimation.py:15-21 (plus --root)."""
    import argparse
    ap = argparse.ArgumentParser(description="mmWave image and surface normals on the GPU")
    ap.add_argument("--name", default="")
    ap.add_argument("--id", default="")
    ap.add_argument("--ext", default="")
    ap.add_argument("--is_los", default="y")
    ap.add_argument("--overwrite_existing", default="n")
    ap.add_argument("--root", default=".", help="folder holding the dataset")
    a = ap.parse_args(argv)
    if not a.id or not a.name:
        raise SystemExit("--id and --name are both needed (no object table here)")
    yes = lambda v: v in ("y", "Y")
    estimate_normals(a.id, a.name, yes(a.is_los), a.ext, yes(a.overwrite_existing),
                     root=a.root)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
