"""The isosurface stage from a stage file (isosurface_optimization.py:30-83
and 379-540): read_sdf_file's reformatting against the stage file itself,
and run_stage against the search and point-cloud functions called directly,
with the reference's two output files."""
import numpy as np
import pytest
import torch

from mmnorm_amd import formats, isosurface, pipeline

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stage_file(tmp_path_factory):
    out = tmp_path_factory.mktemp("stage")
    pipeline.run(cid=2, steps=0, out=str(out))
    return str(out / "sdf_config2.npz")


def test_read_sdf_file_reformats_as_the_reference(stage_file):
    raw = formats.load_sdf_file(stage_file)
    d = isosurface.read_sdf_file(stage_file)
    rsdf = np.asarray(raw["rsdf"], np.float32)
    labels = np.asarray(raw["component_labels"])
    valid = [int(v) for v in raw["valid_labels"] if int(v) != 0]
    sdf = d["sdf"].cpu().numpy()
    assert sdf.shape == rsdf.shape + (len(valid),)
    for i, lab in enumerate(valid):
        inside = labels == lab
        flat = np.where(np.isnan(rsdf), np.float32(isosurface.UNDEF_VAL), rsdf)
        want = np.where(inside, flat, np.float32(isosurface.UNDEF_VAL))
        assert np.array_equal(sdf[..., i], want)
        sel = flat[inside & (flat > isosurface.UNDEF_VAL)]
        assert float(d["tau_bounds"][i, 0]) == float(np.float32(sel.min() + np.float32(0.015)))
        assert float(d["tau_bounds"][i, 1]) == float(np.float32(sel.max() - np.float32(0.015)))
    ch = torch.from_numpy(np.asarray(raw["channels"])) * (0.00001 / 6)
    assert torch.equal(d["channels"].cpu(), ch)
    assert d["rx_dirs"].dtype == torch.float32 and d["num_samples"] == int(raw["num_samples"])


def test_run_stage_equals_the_composed_search(stage_file, tmp_path):
    kw = dict(width=64, height=64, width_fov=20, height_fov=20, num_rx=10, min_dist=0.2,
              max_dist=0.4)
    result, full = isosurface.run_stage(stage_file, str(tmp_path), ext="_t", is_real_data=False,
                                        generator=torch.Generator().manual_seed(3), **kw)
    d = isosurface.read_sdf_file(stage_file)
    taus, grid, losses = isosurface.search_taus(
        d["sdf"], d["rx_locs"], d["rx_dirs"], d["channels"], d["bounding_box"],
        d["wavelengths"], d["voxel_size"], d["tau_bounds"], d["num_samples"],
        is_real_data=False, generator=torch.Generator().manual_seed(3), **kw)
    assert np.array_equal(np.asarray(result["taus"]), np.asarray(taus))
    assert np.array_equal(full["all_losses"], losses.numpy())
    assert np.array_equal(full["all_taus"], grid.numpy())
    sdf = d["sdf"].cpu().numpy()
    pts = np.concatenate([d["coords"][np.abs(sdf[..., k] - taus[k]) < 0.001]
                          for k in range(sdf.shape[-1])])
    assert np.array_equal(result["points"], pts) and len(pts) > 0
    # the reference's two files, with its keys
    r = formats.load_optimization_file(str(tmp_path / "opt_output_optimization_result_t.npz"))
    f = formats.load_optimization_file(str(tmp_path / "opt_output_full_data_t.npz"))
    assert sorted(r) == ["points", "taus"]
    assert sorted(f) == ["all_losses", "all_taus", "coords", "num_samples", "sdf", "sum_img"]
    assert np.array_equal(r["points"], pts)
