"""mmn_rsdf_bfs_noalloc (scratch array in the library, no workspace, no
allocation) against mmn_rsdf_bfs with a caller workspace, through the C ABI:
bit-identical rsdf and hops on two calls in a row, the second on a smaller
grid (the array keeps the first call's contents), and a grid beyond the
array's capacity rejected before anything runs."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from mmnorm_amd import _lib

pytestmark = pytest.mark.gpu

SIX = ndimage.generate_binary_structure(3, 1)


def field(shape, density, seed):
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.normal(size=shape), 1.5)
    mask = noise > np.quantile(noise, 1 - density)
    n = rng.normal(size=shape + (3,))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    n[~mask] = np.nan
    lab, nl = ndimage.label(mask, structure=SIX)
    lab = lab + 1
    seeds = []
    for c in range(2, nl + 2):
        idx = np.flatnonzero((lab == c).ravel())
        seeds.append(idx[len(idx) // 2])
    return n, lab.astype(np.int32), np.array(seeds, np.int32)


def run(entry, n, lab, seeds, dx):
    dev = torch.device("cuda", 0)
    shape = lab.shape
    nrm = torch.from_numpy(n).to(dev)
    labt = torch.from_numpy(lab).to(dev)
    st = torch.from_numpy(seeds).to(dev)
    rsdf = torch.empty(shape, dtype=torch.float64, device=dev)
    hops = torch.empty(shape, dtype=torch.int32, device=dev)
    args = [rsdf.data_ptr(), hops.data_ptr(), nrm.data_ptr(), labt.data_ptr(), st.data_ptr(),
            len(seeds), dx, *shape]
    if entry == "mmn_rsdf_bfs":
        ws_bytes = int(_lib.load().mmn_rsdf_bfs_workspace_size(*shape))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        args += [ws.data_ptr(), ws_bytes]
    _lib.call(entry, *args, _lib.stream_handle(dev))
    torch.cuda.synchronize()
    return rsdf.cpu().numpy(), hops.cpu().numpy()


def test_noalloc_equals_workspace_entry_twice(oracle):
    for shape, density, seed in (((72, 64, 48), 0.45, 21), ((20, 17, 13), 0.6, 22)):
        n, lab, seeds = field(shape, density, seed)
        assert len(seeds) >= 1
        a = run("mmn_rsdf_bfs_noalloc", n, lab, seeds, 1.5e-3)
        b = run("mmn_rsdf_bfs", n, lab, seeds, 1.5e-3)
        assert a[0].tobytes() == b[0].tobytes()
        assert np.array_equal(a[1], b[1])
        w = oracle.bfs_fifo(n, lab, seeds, 1.5e-3)
        assert a[0].tobytes() == w[0].tobytes() and np.array_equal(a[1], w[1])
        assert int((a[1] >= 0).sum()) > 0.2 * np.count_nonzero(~np.isnan(n[..., 0]))


def test_noalloc_rejects_a_grid_beyond_its_scratch():
    # 2048 x 2048 x 4 voxels need about 14 B each: beyond 128 MiB, refused
    # before any HIP call (the pointers are never read)
    with pytest.raises(ValueError, match="scratch"):
        _lib.call("mmn_rsdf_bfs_noalloc", 8, 8, 8, 8, 8, 1, 1e-3, 2048, 2048, 4, None)
