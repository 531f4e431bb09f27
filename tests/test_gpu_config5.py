"""Config 5 past the forward (100x100 aperture at 2 mm, 512 samples 77-81 GHz,
256x256x128 voxels at 1 mm, 20,000 scatterers, antenna positions perturbed
by N(0, 0.3 mm)): the pipeline's front half (pipeline.prepare: forward with
the vote moment, unit normals on the 5 % filter, Li binary image, GPU
labelling and size rule, device start voxels, GPU BFS) against the oracle --
the start voxels and the RSDF and hops bit for bit from the FIFO oracle over
the same fields -- and one optimisation step's position gradient against FP64
central differences."""
import numpy as np
import pytest
import torch

from mmnorm_amd import ops, pipeline

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def state():
    return pipeline.prepare(5, variant="profile")


def test_config5_seeds_rsdf_hops_bit_exact(state, oracle):
    sc = state["sc"]
    nrm = state["normals"].cpu().numpy()
    lab = state["labels"].cpu().numpy()
    m1, m2, m3 = np.meshgrid(sc.y_locs, sc.x_locs, sc.z_locs)
    coords = np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1)
    seeds = oracle.rsdf_seeds(nrm, coords, lab, state["valid"])
    assert len(seeds) >= 1 and np.array_equal(seeds, np.asarray(state["seeds"], np.int64))
    want, whops = oracle.bfs_fifo(nrm, lab, seeds, state["dx"])
    hops = state["hops"].cpu().numpy()
    rsdf = np.ascontiguousarray(state["rsdf"].cpu().numpy())
    assert np.array_equal(hops, whops)
    assert rsdf.view(np.uint64).tobytes() == np.ascontiguousarray(want).view(np.uint64).tobytes()
    # the configuration's field: one large component, hundreds of levels
    assert int((hops >= 0).sum()) > 100_000 and int(hops.max()) > 100


@pytest.mark.parametrize("variant", ["rec", "profile"])
def test_config5_step_gradient_vs_fd(state, oracle, variant):
    """dL/dp at the perturbed start positions of L = sum_{v in S} w |I_v|^2
    (I = backprojection / A) over 64 of the strongest and 64 random voxels,
    against FP64 central differences (h = 1e-6 m) on 12 coordinates; the
    per-coordinate tolerance is 1e-5 of the sum of the magnitudes of the
    coordinate's terms (tests/fd_subset.py)."""
    import fd_subset
    sc, meas, (xl, yl, zl) = state["sc"], state["meas"], state["axes"]
    ant = state["p0"].astype(np.float64)
    flat = state["img"].reshape(-1)
    rng = np.random.default_rng(55)
    top = torch.topk(flat.abs(), 64).indices.cpu().numpy()
    vidx = np.unique(np.concatenate([top, rng.choice(flat.numel(), 64, replace=False)]))
    w = rng.uniform(size=len(vidx))
    scale = 1.0 / len(ant)
    dev = meas.device
    a = torch.from_numpy(state["p0"]).to(dev).requires_grad_(True)
    I = ops.backproject_ad(meas, a, xl, yl, zl, f0=sc.f0, df=sc.df, scale=scale,
                           vox_idx=torch.from_numpy(vidx.astype(np.int32)).to(dev),
                           variant=variant, bbox=state["bbox"])
    (torch.from_numpy(w.astype(np.float32)).to(dev) * I.abs() ** 2).sum().backward()
    grad = a.grad.cpu().numpy().astype(np.float64)
    # rows near the aperture's centre see the strongest voxels: pick coordinates there
    centre = np.argsort(np.linalg.norm(ant[:, :2] - ant[:, :2].mean(0), axis=1))[:400]
    coords = [(int(rng.choice(centre)), int(k)) for k in rng.integers(0, 3, 12)]
    fd, bound = fd_subset.fd_and_bound(oracle, sc, ant, vidx, w, scale, coords)
    got = np.array([grad[i, k] for (i, k) in coords])
    assert np.all(np.abs(got - fd) <= 1e-5 * bound), np.max(np.abs(got - fd) / bound)
    assert np.linalg.norm(got - fd) <= 1e-3 * np.linalg.norm(fd)
