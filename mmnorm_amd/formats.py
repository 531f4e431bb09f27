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
read with ``allow_pickle=False``: every stored field is numeric.  The loader
side of a raw capture -- the ADC files, the trajectory files, the robot-run
records (pickles, read by a restricted unpickler), the experiment JSON and
the dataset's folder scheme -- follows generic_loader.py:36-67,484-630.
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


def save_optimization_file(path, data):
    """Write an isosurface-stage output (generic_loader.py:448-461:
    opt_output<ext>.npz in the processed folder's optimization/)."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    np.savez_compressed(path, **data)


def load_optimization_file(path):
    """The stored dictionary, or None when the file is missing
    (generic_loader.py:463-481)."""
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


# ---------------------------------------------------------------------------
# The dataset's raw captures, as GenericLoader reads them (generic_loader.py).
# ---------------------------------------------------------------------------
def dataset_path(root, obj_id, name, radar_type, is_sim=False, is_los=True, exp_num=1,
                 angles=(0, 0, 0), is_processed=False, is_json=False,
                 path_to_data="sample_data"):
    """The folder GenericLoader._get_path names (generic_loader.py:36-67):
    <root>/<path_to_data>/<id>_<name>/<simulation|robot_collected>/<x>_<y>_<z>/
    <exp>/<los|nlos>/<processed|unprocessed>/radars/<radar_type> (the camera
    folder and the experiment JSON folder as there)."""
    if radar_type is None:
        raise ValueError("You should not use Radar_type = None")
    obj = f"{obj_id}_{name}"
    sim = "simulation" if is_sim else "robot_collected"
    ang = "_".join(str(a) for a in angles)
    los = "" if is_sim or is_los is None else ("los" if is_los else "nlos")
    proc = "processed" if is_processed else "unprocessed"
    if exp_num in ("None", None):
        exp_num = 1
    exp = f"exp{exp_num}" if not is_sim else ""
    base = f"{root}/{path_to_data}/{obj}/{sim}/{ang}"
    if is_json:
        return f"{base}/{exp}"
    if radar_type == "camera":
        return f"{base}/{exp}/{los}/{proc}/camera"
    return f"{base}/{exp}/{los}/{proc}/radars/{radar_type}"


def find_obj_angles(root, obj_id, name, exp_num="2", path_to_data="sample_data"):
    """The object's rotation for a real-data experiment: the angle folder
    (<x>_<y>_<z>) under robot_collected that holds exp<exp_num>
    (generic_loader.py:69-87)."""
    data_folder = f"{root}/{path_to_data}/{obj_id}_{name}/robot_collected"
    for angle_folder in os.listdir(data_folder):
        for exp_folder in os.listdir(f"{data_folder}/{angle_folder}"):
            if exp_folder == f"exp{exp_num}":
                return [int(a) for a in angle_folder.split("_")]
    raise FileNotFoundError(f"Couldn't find any folder matching exp{exp_num} for object "
                            f"{obj_id}_{name}")


def radar_file_timestamp(filename):
    """The capture timestamp in a 77 GHz file name: characters 8-17, or 9-20
    when those are not an integer (generic_loader.py:522-525)."""
    try:
        return int(filename[8:18])
    except ValueError:
        return int(filename[9:21])


def load_radar_files(path, radar_type="77_ghz"):
    """{timestamp: complex (samples, 4)} of every .bin capture in `path`, in
    sorted file-name order (generic_loader.py:484-536, the 77 GHz branch)."""
    if radar_type != "77_ghz":
        raise ValueError("only the 77 GHz captures are read")
    out = {}
    for filename in sorted(os.listdir(path)):
        if filename[-4:] != ".bin":
            continue
        out[radar_file_timestamp(filename)] = read_adc_bin(os.path.join(path, filename))
    return out


def load_robot_loc_files(path):
    """The robot trajectory files of `path` (names starting 'antenna', numeric
    .npz), concatenated key by key in sorted name order (generic_loader.py:
    538-575).  A file whose time and position arrays differ in length is cut,
    as the reference cuts it: once per key of the file, both to the shorter
    length minus 20."""
    combined = {}
    for filename in sorted(f for f in os.listdir(path) if f[:7] == "antenna"):
        with np.load(os.path.join(path, filename), allow_pickle=False) as f:
            cur = {k: f[k] for k in f.files}
        for tag in ("77", "24"):
            tk, lk = f"times_{tag}", f"tx_{tag}_locs"
            if tk in cur and lk in cur and cur[tk].shape[0] != cur[lk].shape[0]:
                for _ in cur:
                    m = min(cur[tk].shape[0], cur[lk].shape[0])
                    cur[tk] = cur[tk][:m - 20]
                    cur[lk] = cur[lk][:m - 20]
        for k, v in cur.items():
            combined[k] = np.concatenate((combined[k], v)) if k in combined else v
    return combined


# Constructors a robot-run record may need: containers, scalars and NumPy
# arrays.  Anything else in a pickle is refused before it is called.
_PICKLE_ALLOWED = {
    ("builtins", n) for n in ("list", "dict", "tuple", "set", "frozenset", "str", "bytes",
                              "bytearray", "int", "float", "complex", "bool", "slice", "range")
} | {
    (m, n) for m in ("numpy.core.multiarray", "numpy._core.multiarray")
    for n in ("_reconstruct", "scalar")
} | {
    (m, "_frombuffer") for m in ("numpy.core.numeric", "numpy._core.numeric")
} | {("numpy", "ndarray"), ("numpy", "dtype"), ("collections", "OrderedDict"),
     ("_codecs", "encode"),  # str -> bytes, how protocol 2 stores an array's data
     ("datetime", "datetime"), ("datetime", "date"), ("datetime", "timedelta")}


def _load_record(f):
    import pickle

    class _Records(pickle.Unpickler):
        def find_class(self, module, name):
            if (module, name) not in _PICKLE_ALLOWED:
                raise pickle.UnpicklingError(
                    f"robot record refers to {module}.{name}, which is not a container, "
                    f"scalar or array constructor")
            return super().find_class(module, name)

    return _Records(f).load()


def load_robot_files(path):
    """The robot-run records of `path` (files named 'continuous*', Python
    pickles of dictionaries), merged key by key in sorted name order by
    concatenation, 'pattern' skipped and then set to 0.5
    (generic_loader.py:577-612).  Unlike the reference's pickle.load, the
    records are read by an unpickler that builds only containers, scalars
    and NumPy arrays and refuses any other callable a file names."""
    robot = {}
    for filename in sorted(f for f in os.listdir(path) if f[:10] == "continuous"):
        with open(os.path.join(path, filename), "rb") as f:
            cur = _load_record(f)
        for k, v in cur.items():
            if k == "pattern":
                continue
            robot[k] = robot[k] + v if k in robot else v
    robot["pattern"] = 0.5
    return robot


def load_exp_json(path):
    """experiment_info.json of an experiment folder (generic_loader.py:614-630)."""
    import json
    with open(os.path.join(path, "experiment_info.json")) as f:
        return json.load(f)


def write_adc_bin(path, samples):
    """Write complex (n, 4) samples as a 77 GHz capture: int16, lanes' I then
    Q per sample (the layout read_adc_bin reads); values are rounded to int16."""
    s = np.asarray(samples)
    raw = np.empty((s.shape[0], 2 * s.shape[1]), "<i2")
    raw[:, :s.shape[1]] = np.clip(np.rint(s.real), -32768, 32767)
    raw[:, s.shape[1]:] = np.clip(np.rint(s.imag), -32768, 32767)
    raw.tofile(path)
