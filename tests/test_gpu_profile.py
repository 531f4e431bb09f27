"""The range-profile forward (mmn_bp_profile) against the FP64 oracle: the
config-1 grid with the vote moment, ragged bistatic TI inputs, voxel lists,
single-sample rows, every row split, config-2 and config-5 full size on a
seeded voxel sample, and its argument checks."""
import os

import numpy as np
import pytest
import torch

from mmnorm_amd import ops, synth

pytestmark = pytest.mark.gpu


def dev():
    return torch.device("cuda", 0)


def T(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev())


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def axes(sc):
    return tuple(T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))


def normal_angles_deg(img, mom, ref, rmom):
    """Vote directions Re(conj(I) m) of the voxels above 5 % of max |ref|."""
    img, ref = img.ravel().astype(np.complex128), ref.ravel()
    strong = np.abs(ref) > 0.05 * np.abs(ref).max()
    m = mom.reshape(-1, 3)[strong].astype(np.complex128)
    rm = rmom.reshape(-1, 3)[strong]
    a = (img[strong].conj()[:, None] * m).real
    b = (ref[strong].conj()[:, None] * rm).real
    c = np.sum(a * b, 1) / np.linalg.norm(a, axis=1) / np.linalg.norm(b, axis=1)
    return np.degrees(np.arccos(np.clip(c, -1, 1)))


def test_config1_grid_with_moment(oracle):
    sc = synth.make_scene(1)
    img, mom = ops.backproject(T(sc.meas, np.complex64), T(sc.ant, np.float32), *axes(sc),
                               f0=sc.f0, df=sc.df, variant="profile", moment=True)
    ref, rmom = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas,
                              np.zeros((1, 3)), sc.f0, sc.df, moment=True)
    img, mom = img.cpu().numpy(), mom.cpu().numpy()
    assert rel(img, ref) <= 1e-5
    assert rel(mom.reshape(-1, 3), rmom) <= 1e-5
    assert normal_angles_deg(img, mom, ref, rmom).max() <= 0.01


@pytest.fixture(scope="module")
def ragged():
    """12x12 TX positions x 3 RX offsets, 77 samples, 12x10x7 voxels (no
    dimension a multiple of the 8x8x8 brick)."""
    cfg = synth.Config(1, 12, 2e-3, 77, (12, 10, 7), 2e-3, 20.0, "points", 5)
    sc = synth.make_scene(1, cfg=cfg, seed=31)
    rxo = np.array([[0.0, 0.0, 0.0], [-3.1e-3, 0.0, 0.0], [0.0, 2.6e-3, 0.0]])
    rng = np.random.default_rng(2)
    rows = np.repeat(sc.ant, 3, axis=0)
    meas = (rng.normal(size=(len(rows), cfg.n_freq)) +
            1j * rng.normal(size=(len(rows), cfg.n_freq))).astype(np.complex64)
    return sc, rows, meas, rxo


@pytest.mark.parametrize("is_ti", [False, True])
@pytest.mark.parametrize("vpt", ["1", "2"])
def test_ragged_bistatic_ti(ragged, oracle, is_ti, vpt):
    sc, rows, meas, rxo = ragged
    os.environ["MMN_PROF_VPT"] = vpt
    try:
        img, mom = ops.backproject(T(meas, np.complex64), T(rows, np.float32), *axes(sc),
                                   T(rxo, np.float32), f0=sc.f0, df=sc.df,
                                   variant="profile", moment=True, is_ti=is_ti)
    finally:
        del os.environ["MMN_PROF_VPT"]
    want, wmom = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df,
                               is_ti=is_ti, moment=True)
    assert rel(img.cpu().numpy(), want) <= 1e-5
    assert rel(mom.cpu().numpy().reshape(-1, 3), wmom) <= 1e-5


def test_voxel_lists(ragged, oracle):
    sc, rows, meas, rxo = ragged
    args = (T(meas, np.complex64), T(rows, np.float32), *axes(sc), T(rxo, np.float32))
    full = ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile").cpu().numpy().ravel()
    for idx in ([417], [3, 3, 3, 839, 3], list(range(0, 840, 7))):
        vi = np.array(idx, np.int32)
        got, gm = ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile",
                                  vox_idx=T(vi, np.int32), moment=True)
        got = got.cpu().numpy()
        want, wm = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0,
                                 sc.df, vidx=vi, moment=True)
        assert rel(got, want) <= 1e-5
        assert rel(gm.cpu().numpy(), wm) <= 1e-5
        assert rel(got, full[vi]) <= 1e-6
        if len(idx) == 5:
            assert got[0] == got[1] == got[2] == got[4]
    empty = ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile",
                            vox_idx=T(np.zeros(0, np.int32), np.int32))
    assert empty.numel() == 0


@pytest.mark.parametrize("F,df", [(1, 15.7e6), (2, 15.7e6), (5, 15.7e6), (3, 0.0),
                                  (33, -62.5e6)])
def test_few_samples_and_zero_step(ragged, oracle, F, df):
    """One sample (the table is constant), two and five samples, df = 0, and
    a descending sweep (df < 0)."""
    sc, rows, _, rxo = ragged
    rng = np.random.default_rng(F)
    meas = (rng.normal(size=(len(rows), F)) + 1j * rng.normal(size=(len(rows), F)))
    meas = meas.astype(np.complex64)
    img = ops.backproject(T(meas, np.complex64), T(rows, np.float32), *axes(sc),
                          T(rxo, np.float32), f0=sc.f0, df=df, variant="profile", is_ti=True)
    ref = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, df,
                        is_ti=True)
    assert rel(img.cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("rs", [1, 3, 8])
def test_every_row_split(ragged, oracle, rs):
    sc, rows, meas, rxo = ragged
    os.environ["MMN_PROF_RS"] = str(rs)
    try:
        img = ops.backproject(T(meas, np.complex64), T(rows, np.float32), *axes(sc),
                              T(rxo, np.float32), f0=sc.f0, df=sc.df, variant="profile")
    finally:
        del os.environ["MMN_PROF_RS"]
    want = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df)
    assert rel(img.cpu().numpy(), want) <= 1e-5


def test_bbox_checks(ragged):
    sc, rows, meas, rxo = ragged
    args = (T(meas, np.complex64), T(rows, np.float32), *axes(sc), T(rxo, np.float32))
    with pytest.raises(ValueError):
        ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile",
                        bbox=[0.0, 0.0, 0.0, -1.0, 1.0, 1.0])
    with pytest.raises(ValueError):
        ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile",
                        bbox=[0.0, 0.0, 0.0, float("nan"), 1.0, 1.0])
    # a box that misses the voxels leaves the table indices clamped: finite output
    out = ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile",
                          bbox=[5.0, 5.0, 5.0, 5.001, 5.001, 5.001])
    assert torch.isfinite(torch.view_as_real(out)).all()


@pytest.mark.parametrize("config,nsample", [(2, 192), (5, 48)])
def test_full_size_sampled_parity(oracle, config, nsample):
    """Full size: the strongest and seeded voxels vs the FP64 oracle, and the
    whole grid vs the recurrence."""
    sc = synth.make_scene(config, device=dev())
    args = (T(sc.meas, np.complex64), T(sc.ant, np.float32), *axes(sc))
    prof = ops.backproject(*args, f0=sc.f0, df=sc.df, variant="profile").reshape(-1)
    rec = ops.backproject(*args, f0=sc.f0, df=sc.df).reshape(-1)
    assert float(torch.linalg.norm(prof - rec) / torch.linalg.norm(rec)) <= 1e-5
    rng = np.random.default_rng(19)
    strong = torch.topk(prof.abs(), nsample).indices.cpu().numpy()
    vidx = np.unique(np.concatenate([strong, rng.choice(sc.cfg.nvox, nsample, replace=False)]))
    ref = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas, np.zeros((1, 3)),
                        sc.f0, sc.df, vidx=vidx)
    got = prof[torch.from_numpy(vidx).to(dev())].cpu().numpy().astype(np.complex128)
    assert rel(got, ref) <= 1e-5
    s = np.isin(vidx, strong)
    assert np.max(np.abs(got[s] - ref[s]) / np.abs(ref[s])) <= 1e-5
