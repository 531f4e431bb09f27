"""The synthetic pipeline end to end on the config-2 scene (60x60 aperture,
256 samples, 128x128x96 voxels at 1.5 mm): forward with the fused vote, Li
binary image, GPU labelling, start voxels from the device fields, GPU BFS and
two optimisation steps through the backward.  The stage file it writes
(sdf.py:226-253 keys) must hold the RSDF that the FIFO oracle computes from
the oracle's own start voxels, bit for bit."""
import numpy as np
import pytest

from mmnorm_amd import formats, pipeline

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", ["rec", "profile"])
def test_pipeline_rsdf_equals_oracle(tmp_path, oracle, variant):
    rec = pipeline.run(cid=2, steps=2, out=str(tmp_path), variant=variant)
    assert rec["components"] >= 1 and rec["reached"] > 0 and rec["band_voxels"] > 0
    assert rec["step_s_mean"] > 0
    d = formats.load_sdf_file(str(tmp_path / "sdf_config2.npz"))
    seeds = oracle.rsdf_seeds(d["normals"], d["coords"], d["component_labels"],
                              d["valid_labels"])
    assert len(seeds) == rec["components"]
    dx = float(d["x_locs"][1] - d["x_locs"][0])
    want, hops = oracle.bfs_fifo(d["normals"], d["component_labels"], seeds, dx)
    assert int((hops >= 0).sum()) == rec["reached"]
    got = np.ascontiguousarray(d["rsdf"], dtype=np.float64)
    assert got.view(np.uint64).tobytes() == np.ascontiguousarray(want).view(np.uint64).tobytes()
