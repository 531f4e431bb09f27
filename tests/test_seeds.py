"""Start voxels of the components (sdf.py:180-191): the device-side search
``sdf.component_seeds_gpu`` (torch on CPU tensors here; the same code runs on
the GPU in the pipeline) against the oracle's ``rsdf_seeds`` on the
reference's coords_full (sdf.py:278-279), bit for bit, ties included."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from mmnorm_amd import sdf

SIX = ndimage.generate_binary_structure(3, 1)


def coords_full(xl, yl, zl):
    m1, m2, m3 = np.meshgrid(yl, xl, zl)
    return np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1)


def axes(shape, pitch):
    return [(np.arange(s) - (s - 1) / 2) * pitch for s in shape]


@pytest.mark.parametrize("shape,density,seed", [((10, 9, 8), 0.4, 1), ((24, 20, 16), 0.3, 2),
                                                ((17, 23, 11), 0.6, 3)])
def test_seeds_equal_oracle(oracle, shape, density, seed):
    rng = np.random.default_rng(seed)
    noise = ndimage.gaussian_filter(rng.normal(size=shape), 1.3)
    mask = noise > np.quantile(noise, 1 - density)
    lab, nl = ndimage.label(mask, structure=SIX)
    lab = (lab + 1).astype(np.int32)
    n = rng.normal(size=shape + (3,))
    n[~mask] = np.nan
    # a few voxels of the components lose their normal (the Li filter)
    n[rng.uniform(size=shape) < 0.1] = np.nan
    xl, yl, zl = axes(shape, 1.5e-3)
    ok = ~np.isnan(n[..., 0])
    valid = [c for c in range(2, nl + 2) if np.any(ok & (lab == c))]
    valid = valid[::-1] + valid[:2]   # any order, with repeats
    want = oracle.rsdf_seeds(n, coords_full(xl, yl, zl), lab, list(dict.fromkeys(valid)))
    got = sdf.component_seeds_gpu(torch.from_numpy(n), torch.from_numpy(lab), valid, xl, yl, zl)
    assert len(want) >= 2
    np.testing.assert_array_equal(got, want)


def test_seed_ties_and_errors(oracle):
    shape = (5, 6, 4)
    n = np.full(shape + (3,), np.nan)
    lab = np.ones(shape, np.int32)
    n[1:3, 1:3, 1:3] = 0.5          # a 2x2x2 cube: eight voxels tie with the mean
    lab[1:3, 1:3, 1:3] = 2
    n[4, 0:4, 0] = 0.1              # a row: even length, two voxels tie
    lab[4, 0:4, 0] = 3
    lab[0, 5, 3] = 4                # a label without normals
    xl, yl, zl = (np.arange(s) * 0.25 for s in shape)
    want = oracle.rsdf_seeds(n, coords_full(xl, yl, zl), lab, [3, 2])
    got = sdf.component_seeds_gpu(torch.from_numpy(n), torch.from_numpy(lab), [3, 2], xl, yl, zl)
    np.testing.assert_array_equal(got, want)
    assert got[1] == np.ravel_multi_index((1, 1, 1), shape)
    assert got[0] == np.ravel_multi_index((4, 1, 0), shape)
    with pytest.raises(ValueError, match="no voxel"):
        sdf.component_seeds_gpu(torch.from_numpy(n), torch.from_numpy(lab), [4], xl, yl, zl)
    with pytest.raises(ValueError, match="axis lengths"):
        sdf.component_seeds_gpu(torch.from_numpy(n), torch.from_numpy(lab), [2], xl[:-1], yl, zl)
    assert sdf.component_seeds_gpu(torch.from_numpy(n), torch.from_numpy(lab), [], xl, yl,
                                   zl).shape == (0,)


def test_seed_labels_far_apart(oracle):
    """Labels more than 2^24 apart take the set-membership path; labels below
    the smallest listed one and above the largest are not selected."""
    shape = (6, 5, 4)
    rng = np.random.default_rng(4)
    n = rng.normal(size=shape + (3,))
    lab = np.zeros(shape, np.int32)
    lab[0:2] = -7
    lab[2:4] = 3
    lab[4:6] = (1 << 25) + 5
    lab[5, 4, 3] = 1 << 30
    xl, yl, zl = axes(shape, 1e-3)
    cf = coords_full(xl, yl, zl)
    for valid in ([3, (1 << 25) + 5], [3], [(1 << 25) + 5, -7]):
        want = oracle.rsdf_seeds(n, cf, lab, valid)
        got = sdf.component_seeds_gpu(torch.from_numpy(n), torch.from_numpy(lab), valid, xl,
                                      yl, zl)
        np.testing.assert_array_equal(got, want)
