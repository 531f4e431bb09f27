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


def test_normal_data_round_trip(tmp_path):
    from mmnorm_amd import formats
    rng = np.random.default_rng(0)
    img = (rng.normal(size=(4, 3, 2)) + 1j * rng.normal(size=(4, 3, 2))).astype(np.complex64)
    axes = [np.arange(n) * 1e-3 for n in (4, 3, 2)]
    ant = rng.normal(size=(10, 3))
    ch = rng.normal(size=(10, 8, 4)) + 0j
    nrm = rng.normal(size=(4, 3, 2, 3))
    p = tmp_path / "normals" / "normal_data.npz"
    formats.save_normal_data(str(p), img, *axes, ant, ch, nrm)
    got = formats.load_normal_data(str(p))
    assert np.array_equal(got[0], img) and all(np.array_equal(a, b) for a, b in zip(got[1], axes))
    assert np.array_equal(got[2], ant) and np.array_equal(got[3], ch)
    assert np.array_equal(got[4], nrm)


def test_sdf_record_round_trip(tmp_path):
    from mmnorm_amd import formats
    rng = np.random.default_rng(1)
    xs, ys, zs = (np.arange(n) * 1e-3 for n in (4, 3, 2))
    ch = rng.normal(size=(5, 16, 4)) + 1j * rng.normal(size=(5, 16, 4))
    rec = formats.sdf_record(np.zeros((4, 3, 2), np.complex64), xs, ys, zs, rng.normal(size=(20, 3)),
                             ch, np.zeros((4, 3, 2, 3)), np.zeros((4, 3, 2)),
                             np.ones((4, 3, 2), np.uint32), [2], np.zeros((4, 3, 2, 3)),
                             77e9, 81e9, 16)
    assert tuple(rec) == formats.SDF_KEYS
    # (N, samples, k) -> (N * k, samples): row 4k + j is antenna k's lane j
    assert np.array_equal(rec["channels"][4 * 2 + 3], ch[2, :, 3])
    assert rec["wavelengths"].shape == (16,)
    assert rec["wavelengths"][0] == pytest.approx(2.99792458e8 / 77e9)
    p = tmp_path / "sdf" / "sdf.npz"
    formats.save_sdf_file(str(p), rec)
    back = formats.load_sdf_file(str(p))
    assert set(back) == set(formats.SDF_KEYS)
    assert np.array_equal(back["channels"], rec["channels"])
    assert formats.load_sdf_file(str(tmp_path / "missing.npz")) is None


def test_read_adc_bin():
    from mmnorm_amd import formats
    vals = np.arange(16, dtype="<i2") - 8        # two samples x (4 I + 4 Q)
    got = formats.read_adc_bin(vals.tobytes())
    assert got.shape == (2, 4)
    assert np.array_equal(got[1], vals[8:12] + 1j * vals[12:16])
    with pytest.raises(ValueError):
        formats.read_adc_bin(vals[:7].tobytes())


def test_correlate_locs_and_meas():
    from mmnorm_amd import image_processor as ip
    lam = 2.99792458e8 / 77e9
    # the robot moves along x at 1 mm per time unit from t = 100 to t = 200,
    # then stands still until t = 300
    t = np.array([100.0, 200.0, 300.0])
    locs = np.array([[0.0, 0.0, 0.0, 7.0], [0.1, 0.0, 0.0, 7.0], [0.1, 0.0, 0.0, 7.0]])
    spc, chirps, frames, sweep, per = 4, 2, 3, 1.0, 1.0
    rd = {1: np.arange(frames * spc * chirps * 4).reshape(-1, 4) + 0j,
          2: np.zeros((frames * spc * chirps, 4)) + 0j,
          3: np.zeros((spc * chirps, 4)) + 0j,     # one frame's worth only
          4: np.zeros((frames * spc * chirps, 4)) + 0j}
    robot = {"all_radar_filenames": ["1", "2", "3", "4", "5"],
             "all_radar_created_ts": [120.0, 250.0, 150.0, 199.0, 160.0]}
    out = ip.correlate_locs_and_meas(robot, rd, t, locs, frames, sweep, per, spc, chirps, 77e9)
    poses = out["poses"]
    # file 1: three frames 2 mm apart (> 3/16 lambda = 0.73 mm), poses interpolated
    assert np.allclose(poses[:3, 0], [0.020, 0.022, 0.024]) and np.all(poses[:3, 3] == 7.0)
    assert np.array_equal(out["radar_data"][1], rd[1][8:12])
    # file 2: the robot stands still -> later frames are duplicates of the first
    assert np.allclose(poses[3, 0], 0.1)
    # file 3: frame 0 kept, later frames lack samples; file 4: frames at
    # t = 199 and 201 are 1 mm apart (kept), the one at 203 repeats 201
    # (dropped); file 5 has no radar file
    assert len(poses) == 3 + 1 + 1 + 2
    assert np.allclose(poses[4, 0], 0.05)
    assert np.allclose(poses[5:, 0], [0.099, 0.1])
    assert 3 / 16 * lam > 2e-4


@pytest.mark.parametrize("shape,use_4_rx,rows", [((5, 7), False, 5), ((5, 7, 1), False, 5),
                                                 ((5, 7, 4), True, 20), ((5, 7, 2), False, 5)])
def test_prep_inputs_reads_rows_as_the_reference_kernel(shape, use_4_rx, rows):
    """The reference flattens the measurements (transposed to (meas, rx,
    samples) when use_4_rx) and its kernel reads row i at i * shape[1]
    (imaging_gpu.cu:100,110): each drop-in row must be that slice."""
    from mmnorm_amd.py_image_gpu import _prep_inputs
    rng = np.random.default_rng(len(shape) * 10 + shape[-1])
    data = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    locs = rng.normal(size=(shape[0], 3))
    rx = [[0.0, 0.0, 0.0]] * (shape[2] if use_4_rx else 1)
    rxo, meas, ant = _prep_inputs(data, rx, locs, use_4_rx)
    S = shape[1]
    flat = (np.transpose(data, (0, 2, 1)) if use_4_rx else data).ravel()
    assert meas.shape == (rows, S) and len(ant) == rows and rxo.shape == (len(rx), 3)
    for i in range(rows):
        np.testing.assert_array_equal(meas[i], flat[i * S:(i + 1) * S])
    if use_4_rx:
        np.testing.assert_array_equal(ant, np.repeat(locs, shape[2], axis=0))
        # row m*R + r holds receiver r's samples of measurement m
        np.testing.assert_array_equal(meas[1 * shape[2] + 3], data[1, :, 3])
    else:
        np.testing.assert_array_equal(meas[2], data.reshape(shape[0], S, -1)[2, :, 0]
                                      if len(shape) == 3 and shape[2] == 1 else meas[2])


def test_prep_inputs_rejects_short_data():
    from mmnorm_amd.py_image_gpu import _prep_inputs
    with pytest.raises(ValueError):
        _prep_inputs(np.zeros((4, 8), np.complex64), [[0, 0, 0]], np.zeros((5, 3)), False)
