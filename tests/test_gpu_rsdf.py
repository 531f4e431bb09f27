"""GPU RSDF breadth-first search: bit-exact against the FIFO restatement."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from mmnorm_amd import sdf as msdf

pytestmark = pytest.mark.gpu

SIX = ndimage.generate_binary_structure(3, 1)


def random_field(shape, density, seed, blobs=True):
    rng = np.random.default_rng(seed)
    if blobs:
        noise = ndimage.gaussian_filter(rng.normal(size=shape), 1.5)
        mask = noise > np.quantile(noise, 1 - density)
    else:
        mask = rng.uniform(size=shape) < density
    n = rng.normal(size=shape + (3,))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    n[~mask] = np.nan
    labeled, nlab = ndimage.label(mask, structure=SIX)
    labeled = labeled + 1  # background 1, components 2.. as cc3d + 1
    return n, labeled, nlab


def axes_of(shape, pitch):
    return [(np.arange(s) - (s - 1) / 2) * pitch for s in shape]


def coords_of(shape, pitch):
    xl, yl, zl = axes_of(shape, pitch)
    m1, m2, m3 = np.meshgrid(yl, xl, zl)
    return np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1), xl


@pytest.mark.parametrize("shape,density,seed,blobs", [
    ((12, 10, 9), 0.5, 1, True),
    ((20, 24, 16), 0.35, 2, True),
    ((16, 16, 16), 0.6, 3, False),
    ((40, 33, 27), 0.3, 4, True),
    # frontiers above one workgroup's 4096 entries (single seed, peak 9225): the
    # single-workgroup and multi-block level paths hand over both ways
    ((96, 96, 64), 0.9, 5, False),
    # 8387 seeds: the first level already runs on the multi-block path
    ((64, 64, 64), 0.25, 6, False),
])
def test_bfs_bit_exact_vs_fifo(shape, density, seed, blobs, oracle):
    n, labeled, nlab = random_field(shape, density, seed, blobs)
    coords, xl = coords_of(shape, 1.5e-3)
    sizes = ndimage.sum(np.ones(shape), labeled, np.arange(2, nlab + 2))
    valid = np.arange(2, nlab + 2)[sizes >= 2]
    seeds = msdf.component_seeds(n, coords, labeled, valid)
    np.testing.assert_array_equal(seeds, oracle.rsdf_seeds(n, coords, labeled, valid))
    # the device-side seed search on the GPU's copies of the fields
    dev = torch.device("cuda", 0)
    seeds_dev = msdf.component_seeds_gpu(
        torch.from_numpy(n).to(dev), torch.from_numpy(labeled.astype(np.int32)).to(dev),
        valid, *axes_of(shape, 1.5e-3))
    np.testing.assert_array_equal(seeds_dev, seeds)
    got, hops = msdf.rsdf_components(n, labeled, seeds, xl[1] - xl[0], return_hops=True)
    want, whops = oracle.bfs_fifo(n, labeled, seeds, xl[1] - xl[0])
    assert got.view(np.uint64).tobytes() == want.view(np.uint64).tobytes()
    np.testing.assert_array_equal(hops, whops)


@pytest.mark.parametrize("env", [
    {"MMN_BFS_FUSED": "0"},
    {"MMN_BFS_FUSED": "0", "MMN_BFS_GRID": "1024"},
    {"MMN_BFS_PER": "2"},
    {"MMN_BFS_PER": "4", "MMN_BFS_CAP": "4096"},
    {"MMN_BFS_PER": "8", "MMN_BFS_CAP": "8192"},
    {"MMN_BFS_PER": "0"},
    {"MMN_BFS_CAP": "16", "MMN_BFS_BATCH": "1"},
    {"MMN_BFS_FGRID": "16"},
    {"MMN_BFS_FGRID": "1024", "MMN_BFS_BATCH": "256"},
], ids=lambda e: ",".join(f"{k[8:]}={v}" for k, v in e.items()))
def test_bfs_launch_shapes_bit_exact(env, oracle):
    """The BFS's other launch shapes (the two-kernel levels, positions per
    thread of the single-workgroup loop, its frontier cap, levels per batch,
    grid sizes) on the single-seed field whose frontier peaks at 9,225
    entries, so that both level forms run and hand over both ways: bit-exact
    against the C FIFO restatement."""
    import os
    shape = (96, 96, 64)
    n, labeled, nlab = random_field(shape, 0.9, 5, False)
    coords, xl = coords_of(shape, 1.5e-3)
    sizes = ndimage.sum(np.ones(shape), labeled, np.arange(2, nlab + 2))
    valid = np.arange(2, nlab + 2)[sizes >= 2]
    seeds = oracle.rsdf_seeds(n, coords, labeled, valid)
    want, whops = oracle.bfs_fifo(n, labeled, seeds, xl[1] - xl[0])
    os.environ.update(env)
    try:
        got, hops = msdf.rsdf_components(n, labeled, seeds, xl[1] - xl[0], return_hops=True)
    finally:
        for k in env:
            del os.environ[k]
    assert got.view(np.uint64).tobytes() == want.view(np.uint64).tobytes()
    np.testing.assert_array_equal(hops, whops)


def test_bfs_matches_python_fifo_small(oracle):
    n, labeled, nlab = random_field((7, 6, 5), 0.55, 9, False)
    coords, xl = coords_of((7, 6, 5), 1e-3)
    valid = np.arange(2, nlab + 2)
    seeds = msdf.component_seeds(n, coords, labeled, valid)
    got, hops = msdf.rsdf_components(n, labeled, seeds, xl[1] - xl[0], return_hops=True)
    want, whops = oracle.py_bfs_fifo(n, labeled, seeds, xl[1] - xl[0])
    assert np.array_equal(got, want, equal_nan=True)
    np.testing.assert_array_equal(hops, whops)


def test_constant_normal_closed_form():
    """Normals (0,0,1) on a full box: rsdf = dx * (k - k0), hops = Manhattan distance."""
    shape = (9, 8, 11)
    n = np.zeros(shape + (3,))
    n[..., 2] = 1.0
    labeled = np.full(shape, 2)
    dx = 1e-3
    seed = np.ravel_multi_index((4, 3, 5), shape)
    rsdf, hops = msdf.rsdf_components(n, labeled, [seed], dx, return_hops=True)
    i, j, k = np.indices(shape)
    assert np.allclose(rsdf, dx * (k - 5), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(hops, np.abs(i - 4) + np.abs(j - 3) + np.abs(k - 5))


def test_drop_in_construct_rsdf_components(oracle):
    n, labeled, nlab = random_field((18, 15, 12), 0.45, 21, True)
    coords, xl = coords_of((18, 15, 12), 1e-3)
    valid = np.arange(2, nlab + 2)[::2]
    got = msdf._construct_rsdf_components(n, coords, labeled, valid, xl)
    seeds = oracle.rsdf_seeds(n, coords, labeled, valid)
    want, _ = oracle.bfs_fifo(n, labeled, seeds, xl[1] - xl[0])
    assert got.view(np.uint64).tobytes() == want.view(np.uint64).tobytes()


def test_bfs_deterministic_and_empty():
    n, labeled, nlab = random_field((30, 30, 20), 0.4, 5, True)
    coords, xl = coords_of((30, 30, 20), 1e-3)
    valid = np.arange(2, nlab + 2)
    a = msdf._construct_rsdf_components(n, coords, labeled, valid, xl)
    b = msdf._construct_rsdf_components(n, coords, labeled, valid, xl)
    assert a.tobytes() == b.tobytes()
    e = msdf._construct_rsdf_components(n, coords, labeled, [], xl)
    assert np.all(np.isnan(e))


def test_bfs_rejects_bad_seed():
    n = np.zeros((3, 3, 3, 3))
    with pytest.raises(ValueError):
        msdf.rsdf_components(n, np.ones((3, 3, 3), np.int32), [27], 1e-3)
    assert torch.cuda.is_available()


@pytest.mark.parametrize("n_seeds", [3, 1500])
def test_c_abi_rejects_a_seed_outside_the_grid(n_seeds):
    """mmn_rsdf_bfs reads the seeds on the device: one outside [0, V) is
    skipped there, no level runs, and the call returns MMN_ERR_INPUT.  1500
    seeds start on the multi-block level kernels, 3 on the single workgroup.
    A valid call on the same workspace afterwards is unaffected."""
    from mmnorm_amd import _lib
    dev = torch.device("cuda", 0)
    shape = (16, 12, 10)
    n, lab, _ = random_field(shape, 0.9, 11, blobs=False)
    V = int(np.prod(shape))
    nrm = torch.from_numpy(n).to(dev)
    labt = torch.from_numpy(lab.astype(np.int32)).to(dev)
    good = np.flatnonzero(~np.isnan(n[..., 0]).ravel())[:n_seeds].astype(np.int32)
    for bad in (V, -1):
        seeds = good.copy()
        seeds[len(seeds) // 2] = bad
        st = torch.from_numpy(seeds).to(dev)
        rsdf = torch.empty(shape, dtype=torch.float64, device=dev)
        hops = torch.empty(shape, dtype=torch.int32, device=dev)
        lib = _lib.load()
        ws_bytes = int(lib.mmn_rsdf_bfs_workspace_size(*shape))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        with pytest.raises(ValueError, match="outside the grid"):
            _lib.call("mmn_rsdf_bfs", rsdf.data_ptr(), hops.data_ptr(), nrm.data_ptr(),
                      labt.data_ptr(), st.data_ptr(), len(seeds), 1e-3, *shape, ws.data_ptr(),
                      ws_bytes, _lib.stream_handle(dev))
    # the same workspace serves a valid call
    st = torch.from_numpy(good[:1]).to(dev)
    _lib.call("mmn_rsdf_bfs", rsdf.data_ptr(), hops.data_ptr(), nrm.data_ptr(), labt.data_ptr(),
              st.data_ptr(), 1, 1e-3, *shape, ws.data_ptr(), ws_bytes, _lib.stream_handle(dev))
    torch.cuda.synchronize()
    assert int(hops.max()) > 0
