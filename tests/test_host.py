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
    """ImageProcessor.correlate_locs_and_meas with the reference's argument
    list and loader-shaped dicts (image_processor.py:74-154)."""
    from mmnorm_amd import image_processor as ip
    lam = 2.99792458e8 / 77e9
    # the robot moves along x at 1 mm per time unit from t = 100 to t = 200,
    # then stands still until t = 300
    t = np.array([100.0, 200.0, 300.0])
    locs = np.array([[0.0, 0.0, 0.0, 7.0], [0.1, 0.0, 0.0, 7.0], [0.1, 0.0, 0.0, 7.0]])
    spc, chirps, frames, sweep, per = 4, 2, 3, 1.0, 1.0
    params = ip.params_with(num_frames=frames, sweep_time=sweep, periodicity=per,
                            num_samples=spc, num_chirps=chirps, min_f=77e9)
    rd = {1: np.arange(frames * spc * chirps * 4).reshape(-1, 4) + 0j,
          2: np.zeros((frames * spc * chirps, 4)) + 0j,
          3: np.zeros((spc * chirps, 4)) + 0j,     # one frame's worth only
          4: np.zeros((frames * spc * chirps, 4)) + 0j,
          6: np.zeros((frames * spc * chirps, 4)) + 0j}
    robot = {"all_radar_filenames": ["1", "2", "3", "4", "5", "6", "7"],
             "all_radar_created_ts": [120.0, 250.0, 150.0, 199.0, 160.0, 99.5, -1]}
    proc = ip.ImageProcessor(False, params=params)
    out = proc.correlate_locs_and_meas("77_ghz", robot, rd, {"aperture size": "almost-square"},
                                       {"times_77": t, "tx_77_locs": locs}, speed="speed_8")
    poses = out["poses"]
    # file 1: three frames 2 mm apart (> 3/16 lambda = 0.73 mm), poses interpolated
    assert np.allclose(poses[:3, 0], [0.020, 0.022, 0.024]) and np.all(poses[:3, 3] == 7.0)
    assert np.array_equal(out["radar_data"][1], rd[1][8:12])
    # file 2: the robot stands still -> later frames are duplicates of the first
    assert np.allclose(poses[3, 0], 0.1)
    # file 3: frame 0 kept, later frames lack samples; file 4: frames at
    # t = 199 and 201 are 1 mm apart (kept), the one at 203 repeats 201
    # (dropped); file 5 has no radar file; file 6 starts before the
    # trajectory (frame 0 dropped) so its later frames are dropped too, as in
    # the reference; file 7 has no creation time
    assert len(poses) == 3 + 1 + 1 + 2
    assert np.allclose(poses[4, 0], 0.05)
    assert np.allclose(poses[5:, 0], [0.099, 0.1])
    assert 3 / 16 * lam > 2e-4
    # the motion-capture frame swaps axes (image_processor.py:64-72)
    opt = ip.ImageProcessor(True, params=params).correlate_locs_and_meas(
        "77_ghz", robot, rd, {"aperture size": "almost-square"},
        {"times_77": t, "tx_77_locs": locs[:, :3]})
    assert np.allclose(opt["poses"][0], [-0.020, 0.0, 0.0])
    with pytest.raises(AssertionError):
        proc.correlate_locs_and_meas("10_ghz", robot, rd, {"aperture size": "almost-square"},
                                     {"times_77": t, "tx_77_locs": locs})


def test_get_radar_parameters():
    from mmnorm_amd import image_processor as ip
    p = ip.get_radar_parameters("77_ghz", False, "almost-square")
    assert p["min_f"] == 77.5001e9 and p["num_samples"] == 512 and p["slope"] == 5861160
    # slope * samples is the bandwidth the chirp sweeps (image_processor.py:359)
    assert abs(ip.SLOPE_REAL * 512 - p["bandwidth"]) < 1e-5 * p["bandwidth"]
    # 'large' looks up 'normal', which the reference's table lacks for 77 GHz
    with pytest.raises(KeyError):
        ip.get_radar_parameters("77_ghz", False, "large")
    with pytest.raises(KeyError):
        ip.get_radar_parameters("77_ghz", True)
    assert ip.ImageProcessor().use_interpolated_processing is True


def test_dataset_path_and_capture_names(tmp_path):
    from mmnorm_amd import formats
    got = formats.dataset_path("/r", "001", "cup", "77_ghz", is_los=False, exp_num="2",
                               angles=(0, 90, 0))
    assert got == "/r/sample_data/001_cup/robot_collected/0_90_0/exp2/nlos/unprocessed/radars/77_ghz"
    assert formats.dataset_path("/r", "001", "cup", "77_ghz", exp_num=None, is_json=True) == \
        "/r/sample_data/001_cup/robot_collected/0_0_0/exp1"
    assert formats.radar_file_timestamp("adc_data1712345678.bin") == 1712345678
    assert formats.radar_file_timestamp("adc_data_171234567890.bin") == 171234567890
    # captures round-trip through the int16 I/Q layout, keyed by timestamp
    rng = np.random.default_rng(1)
    a = np.rint(rng.normal(size=(6, 4)) * 900) + 1j * np.rint(rng.normal(size=(6, 4)) * 900)
    formats.write_adc_bin(str(tmp_path / "adc_data1700000001.bin"), a)
    formats.write_adc_bin(str(tmp_path / "adc_data1700000002.bin"), 2 * a)
    (tmp_path / "notes.txt").write_text("x")
    rd = formats.load_radar_files(str(tmp_path))
    assert sorted(rd) == [1700000001, 1700000002]
    assert np.array_equal(rd[1700000001], a) and np.array_equal(rd[1700000002], 2 * a)
    raw = np.fromfile(str(tmp_path / "adc_data1700000001.bin"), "<i2").reshape(6, 8)
    assert np.array_equal(raw[:, :4], a.real) and np.array_equal(raw[:, 4:], a.imag)


def test_robot_loc_files_trim_rule(tmp_path):
    """A trajectory file whose times and positions differ in length is cut to
    the shorter length minus 20, once per key of the file (generic_loader.py:
    554-563): four keys, so 80 samples go."""
    from mmnorm_amd import formats
    np.savez(tmp_path / "antenna_1.npz", times_77=np.arange(200.0),
             tx_77_locs=np.ones((205, 3)), times_24=np.arange(10.0), tx_24_locs=np.ones((10, 3)))
    np.savez(tmp_path / "antenna_2.npz", times_77=np.arange(200.0, 210.0),
             tx_77_locs=np.zeros((10, 3)), times_24=np.arange(3.0), tx_24_locs=np.ones((3, 3)))
    np.savez(tmp_path / "other.npz", times_77=np.arange(3.0))
    c = formats.load_robot_loc_files(str(tmp_path))
    assert c["times_77"].shape == (120 + 10,) and c["tx_77_locs"].shape == (130, 3)
    assert c["times_77"][119] == 119.0 and c["times_77"][120] == 200.0
    assert c["times_24"].shape == (13,)
    import json
    (tmp_path / "experiment_info.json").write_text(json.dumps({"aperture size": "almost-square"}))
    assert formats.load_exp_json(str(tmp_path))["aperture size"] == "almost-square"


class _Planted:
    """Pickles as a call of os.getcwd: a record that names a callable."""

    def __reduce__(self):
        import os
        return (os.getcwd, ())


def test_robot_files_merge_and_refuse_callables(tmp_path):
    """generic_loader.py:577-612: 'continuous*' records merged key by key in
    sorted name order (lists concatenated), 'pattern' replaced by 0.5; other
    files ignored.  A record naming a callable is refused unread."""
    import pickle
    from mmnorm_amd import formats
    with open(tmp_path / "continuous_b.pkl", "wb") as f:
        pickle.dump({"all_radar_filenames": ["3", "4"], "all_radar_created_ts": [3.0, 4.0],
                     "pattern": 2, "extra": np.arange(3)}, f)
    with open(tmp_path / "continuous_b2.pkl", "wb") as f:
        pickle.dump({"extra": np.arange(3, 6.0)}, f, protocol=2)
    with open(tmp_path / "continuous_a.pkl", "wb") as f:
        pickle.dump({"all_radar_filenames": ["1"], "all_radar_created_ts": [1.0],
                     "pattern": 7}, f, protocol=5)
    with open(tmp_path / "simple_0.pkl", "wb") as f:
        pickle.dump({"all_radar_filenames": ["x"]}, f)
    r = formats.load_robot_files(str(tmp_path))
    assert r["all_radar_filenames"] == ["1", "3", "4"]
    assert r["all_radar_created_ts"] == [1.0, 3.0, 4.0]
    # arrays merge with +, as the reference's robot_data[key] + cur_file[key]
    assert r["pattern"] == 0.5 and np.array_equal(r["extra"], np.arange(3) + np.arange(3, 6.0))
    with open(tmp_path / "continuous_c.pkl", "wb") as f:
        pickle.dump({"all_radar_filenames": [_Planted()]}, f)
    with pytest.raises(pickle.UnpicklingError, match="posix.getcwd|os.getcwd"):
        formats.load_robot_files(str(tmp_path))


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


def _rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return rz @ ry @ rx


def test_hull_face_box_rotated_boxes():
    """A box's hull faces are its faces: the volume is the box's, however it
    is turned, with random points inside that do not change the hull."""
    rng = np.random.default_rng(3)
    for _ in range(5):
        ext = rng.uniform(0.01, 0.2, 3)
        corners = np.stack(np.meshgrid(*[[0.0, e] for e in ext], indexing="ij"), -1).reshape(-1, 3)
        inner = rng.uniform(0, 1, (500, 3)) * ext
        r = _rot(*rng.uniform(0, 2 * np.pi, 3))
        pts = np.concatenate([corners, inner]) @ r.T + rng.normal(size=3)
        assert sdf.hull_face_box_volume(pts) == pytest.approx(np.prod(ext), rel=1e-9)
    g = np.stack(np.meshgrid(np.arange(10), np.arange(6), np.arange(3), indexing="ij"),
                 -1).reshape(-1, 3) * 1e-3
    assert sdf.hull_face_box_volume(g) == pytest.approx(9e-3 * 5e-3 * 2e-3, rel=1e-9)


def test_hull_face_box_regular_tetrahedron():
    """Open3d's face-framed boxes, not the true minimum: for a regular
    tetrahedron of edge s every triangle frame gives s (an edge) x s sqrt(3)/2
    (the face's height) x s sqrt(2/3) (the solid's height) = s^3 / sqrt 2,
    while the true minimal box (the cube the tetrahedron is inscribed in) is
    s^3 / (2 sqrt 2); the principal-axes box of this set is neither."""
    s = 0.1
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) * s / (2 * np.sqrt(2))
    r = _rot(0.3, 1.1, -0.4)
    v = sdf.hull_face_box_volume(tet @ r.T)
    assert v == pytest.approx(s ** 3 / np.sqrt(2), rel=1e-9)
    assert v > 1.9 * s ** 3 / (2 * np.sqrt(2))


def test_hull_face_box_degenerate_sets():
    assert sdf.hull_face_box_volume(np.zeros((3, 3))) == 0.0
    flat = np.random.default_rng(1).uniform(size=(50, 3))
    flat[:, 2] = 0.5
    assert sdf.hull_face_box_volume(flat) == 0.0


def test_find_obj_angles(tmp_path):
    """generic_loader.py:69-87: the angle folder under robot_collected that
    holds exp<n>; none: an error."""
    from mmnorm_amd import formats
    base = tmp_path / "sample_data" / "048_hammer" / "robot_collected"
    (base / "0_0_0" / "exp1").mkdir(parents=True)
    (base / "0_90_-45" / "exp2").mkdir(parents=True)
    assert formats.find_obj_angles(str(tmp_path), "048", "hammer", "2") == [0, 90, -45]
    assert formats.find_obj_angles(str(tmp_path), "048", "hammer", "1") == [0, 0, 0]
    with pytest.raises(FileNotFoundError):
        formats.find_obj_angles(str(tmp_path), "048", "hammer", "3")
