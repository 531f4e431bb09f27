"""Stage files of the reference pipeline and the raw ADC reader, so that runs
of this package and of the reference can hand data to each other:

* ``normal_data.npz`` (generic_loader.py:357-407, written after the normal
  estimation): sum_img, x/y/z_locs, antenna_locs, channels, normals;
* ``sdf.npz`` (sdf.py:226-253 and generic_loader.py:413-446, written after
  the RSDF): the keys of ``sdf_record``;
* the 77 GHz ADC capture (generic_loader.py:521-534): little-endian int16,
  eight values per sample -- four lanes' in-phase then four lanes'
  quadrature parts.

Files are written with ``numpy.savez_compressed`` as the reference does and
read with ``allow_pickle=False``: every stored field is numeric.
"""
from __future__ import annotations

import os

import numpy as np

C0 = 2.99792458e8
NORMAL_KEYS = ("sum_img", "x_locs", "y_locs", "z_locs", "antenna_locs", "channels", "normals")
SDF_KEYS = ("rsdf", "sum_img", "normals", "rx_locs", "rx_dirs", "channels", "component_labels",
            "valid_labels", "coords", "x_locs", "y_locs", "z_locs", "num_samples",
            "bounding_box", "wavelengths", "voxel_size")


def save_normal_data(path, sum_img, x_locs, y_locs, z_locs, antenna_locs, channels, normals):
    """Write normal_data.npz with the keys of generic_loader.py:377-382."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    np.savez_compressed(path, sum_img=sum_img, x_locs=x_locs, y_locs=y_locs, z_locs=z_locs,
                        antenna_locs=np.array(antenna_locs), channels=channels, normals=normals)


def load_normal_data(path):
    """(sum_img, (x_locs, y_locs, z_locs), antenna_locs, channels, normals), the
    return of generic_loader.load_surface_normals."""
    with np.load(path, allow_pickle=False) as f:
        return (f["sum_img"], (f["x_locs"], f["y_locs"], f["z_locs"]), f["antenna_locs"],
                f["channels"], f["normals"])


def sdf_record(sum_img, x_locs, y_locs, z_locs, antenna_locs, channels, all_normals, rsdf,
               labeled, valid_labels, coords_full, min_f, max_f, num_samples):
    """The dictionary sdf.py:_save_sdf_file writes (sdf.py:230-252): channels
    (N, num_samples, k) reordered to (N * k, num_samples), boresights +z, the
    chirp's wavelengths, the grid's bounding box and voxel size."""
    ch = np.asarray(channels)
    if ch.ndim == 3:
        ch = np.transpose(ch, (0, 2, 1))
        ch = ch.reshape(ch.shape[0] * ch.shape[1], ch.shape[2])
    ant = np.asarray(antenna_locs)
    rx_dirs = np.zeros_like(ant)
    rx_dirs[:, -1] = 1.0
    return {
        "rsdf": rsdf, "sum_img": sum_img, "normals": all_normals, "rx_locs": ant,
        "rx_dirs": rx_dirs, "channels": ch, "component_labels": labeled,
        "valid_labels": np.asarray(valid_labels), "coords": coords_full,
        "x_locs": x_locs, "y_locs": y_locs, "z_locs": z_locs, "num_samples": num_samples,
        "bounding_box": np.array([x_locs[0], y_locs[0], z_locs[0], x_locs[-1], y_locs[-1],
                                  z_locs[-1]]),
        "wavelengths": C0 / np.arange(min_f, max_f, (max_f - min_f) / num_samples),
        "voxel_size": z_locs[1] - z_locs[0],
    }


def save_sdf_file(path, data):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    np.savez_compressed(path, **data)


def load_sdf_file(path):
    """The stored dictionary (arrays), or None when the file is missing, as
    generic_loader.load_sdf_file returns."""
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def read_adc_bin(src, num_lanes=4):
    """Complex ADC samples (n, num_lanes) from a 77 GHz capture: a path, bytes
    or an int16 array; row r holds lanes' I in values 0..3 and Q in 4..7."""
    if isinstance(src, (bytes, bytearray, memoryview)):
        raw = np.frombuffer(src, dtype="<i2")
    elif isinstance(src, np.ndarray):
        raw = src.astype("<i2", copy=False).ravel()
    else:
        raw = np.fromfile(src, dtype="<i2")
    per = num_lanes * 2
    if raw.size % per:
        raise ValueError(f"{raw.size} values is not a multiple of {per}")
    rows = raw.reshape(raw.size // per, per)
    return rows[:, :num_lanes] + 1j * rows[:, num_lanes:]
