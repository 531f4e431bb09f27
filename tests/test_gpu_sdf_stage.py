"""The RSDF stage from its stage file (sdf.py:255-300): normal_data.npz of
the config-2 scene in, sdf.npz out, equal to the drop-ins composed by hand
and to the FIFO oracle's RSDF from the oracle's own start voxels, bit for
bit; and the folder-scheme entry point with its skip rule."""
import os

import numpy as np
import pytest
import torch

from mmnorm_amd import formats, image_processor, ops, sdf, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def normal_file(tmp_path_factory):
    dev = torch.device("cuda", 0)
    sc = synth.make_scene(2)
    T = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(dev)
    img, mom = ops.backproject(T(sc.meas, np.complex64), T(sc.ant, np.float32),
                               T(sc.x_locs, np.float32), T(sc.y_locs, np.float32),
                               T(sc.z_locs, np.float32), f0=sc.f0, df=sc.df, moment=True)
    img = img / sc.meas.shape[0]
    mag = img.abs().reshape(-1)
    vidx = torch.nonzero(mag > 0.05 * mag.max()).reshape(-1).to(torch.int32)
    nrm = ops.sn_from_moment(img.reshape(-1), mom.reshape(-1, 3)[vidx.long()].contiguous(),
                             vidx).double()
    normals = torch.full((img.numel(), 3), float("nan"), dtype=torch.float64, device=dev)
    normals[vidx.long()] = nrm
    path = str(tmp_path_factory.mktemp("normals") / "normal_data.npz")
    formats.save_normal_data(path, img.cpu().numpy(), sc.x_locs, sc.y_locs, sc.z_locs, sc.ant,
                             sc.meas[:, :, None], normals.reshape(*img.shape, 3).cpu().numpy())
    return path


def test_compute_rsdf_file_equals_composed_dropins_and_oracle(normal_file, tmp_path, oracle):
    out = str(tmp_path / "sdf.npz")
    data = sdf.compute_rsdf_file(normal_file, out)
    d = formats.load_sdf_file(out)
    assert sorted(d) == sorted(formats.SDF_KEYS)
    # the same steps by hand
    sum_img, (xl, yl, zl), ant, ch, normals = formats.load_normal_data(normal_file)
    m1, m2, m3 = np.meshgrid(yl, xl, zl)
    coords = np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1)
    filt = sdf._separate_binary_image(sum_img, coords)
    n = np.array(normals, copy=True)
    n[filt] = np.nan
    n /= np.linalg.norm(n, axis=-1)[..., None].repeat(3, axis=-1)
    labeled, valid = sdf._find_connected_components(n, xl)
    assert np.array_equal(d["component_labels"], labeled)
    assert np.array_equal(d["valid_labels"], np.asarray(valid)) and len(valid) > 0
    assert np.array_equal(d["normals"], n, equal_nan=True)
    # the FIFO oracle from its own start voxels, bit for bit
    seeds = oracle.rsdf_seeds(n, coords, labeled, valid)
    want, hops = oracle.bfs_fifo(n, labeled, seeds, float(xl[1] - xl[0]))
    got = np.ascontiguousarray(d["rsdf"], dtype=np.float64)
    assert got.view(np.uint64).tobytes() == np.ascontiguousarray(want).view(np.uint64).tobytes()
    assert int((hops >= 0).sum()) > 1000
    # the chirp of the 77 GHz almost-square rig
    p = image_processor.get_radar_parameters("77_ghz", False, "almost-square")
    assert int(d["num_samples"]) == p["num_samples"] and len(d["wavelengths"]) == p["num_samples"]
    assert np.array_equal(d["rsdf"], data["rsdf"], equal_nan=True)


def test_compute_rsdf_folder_scheme_and_skip(normal_file, tmp_path, capsys):
    root = str(tmp_path)
    proc = formats.dataset_path(root, "048", "hammer", "77_ghz", is_los=True, exp_num="2",
                                is_processed=True)
    os.makedirs(f"{proc}/normals")
    with open(normal_file, "rb") as src, open(f"{proc}/normals/normal_data_x.npz", "wb") as dst:
        dst.write(src.read())
    assert sdf.main(["--root", root, "--ext", "_x"]) == 0
    assert os.path.exists(f"{proc}/sdf/sdf_x.npz")
    assert sdf.compute_rsdf("048", "hammer", True, "_x", False, root=root) is None
    assert "skipping" in capsys.readouterr().out
