"""Host-side logic of the drop-ins that needs no GPU: the size rule and the
stand-in volume of the binary-image test."""
import numpy as np
import pytest

from mmnorm_amd import sdf


def test_pca_box_volume_axis_aligned_and_rotated():
    g = np.stack(np.meshgrid(np.arange(10), np.arange(6), np.arange(3), indexing="ij"),
                 -1).reshape(-1, 3) * 1e-3
    assert sdf.pca_box_volume(g) == pytest.approx(9e-3 * 5e-3 * 2e-3, rel=1e-9)
    # a rigid rotation does not change it
    th = 0.7
    rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    assert sdf.pca_box_volume(g @ rot.T + 0.3) == pytest.approx(9e-3 * 5e-3 * 2e-3, rel=1e-9)
    assert sdf.pca_box_volume(g[:1]) == 0.0


def test_size_rule_matches_the_reference_rule():
    """sdf.py:92-119 written out again for 1 mm voxels: > 30*30*40 voxels,
    else > 15*15*20."""
    dx = 1e-3
    areas = np.array([0, 5, 36001, 36000, 0], float)
    assert list(sdf.size_rule(areas, dx)) == [3]
    areas = np.array([0, 4501, 4500, 10], float)
    assert list(sdf.size_rule(areas, dx)) == [2]
    assert len(sdf.size_rule(np.array([0, 10], float), dx)) == 0


def test_select_tau_one_component():
    from mmnorm_amd import isosurface
    taus = np.linspace(-0.02, 0.02, 41)[:, None]
    # two dips; the deeper one first, a nearly as deep one later (closer to
    # the top): the later peak of the flipped loss is chosen
    x = np.arange(41)
    loss = 50 - 30 * np.exp(-((x - 10) / 2.0) ** 2) - 28 * np.exp(-((x - 30) / 2.0) ** 2)
    assert isosurface._select_tau(taus, loss)[0] == pytest.approx(taus[30, 0])
    # a single dip
    loss = 50 - 30 * np.exp(-((x - 12) / 2.0) ** 2)
    assert isosurface._select_tau(taus, loss)[0] == pytest.approx(taus[12, 0])


def test_range_window_bins():
    from mmnorm_amd import isosurface
    f = 77e9 + np.arange(512) * (4e9 / 512)
    isosurface.set_range_window(512, 2.99792458e8 / f)
    d = isosurface.DISTANCES
    assert isosurface.MIN_DIST_IDX == np.count_nonzero(d < 0.35)
    assert isosurface.MAX_DIST_IDX == np.count_nonzero(d < 0.55)
    assert 0 < isosurface.MIN_DIST_IDX < isosurface.MAX_DIST_IDX < 512
