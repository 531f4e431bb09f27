"""Edge cases of the recurrence forward against the FP64 oracle: sample counts
that leave a short last block, several RX offsets with the TI range offset,
one-voxel and empty voxel lists, repeated voxels, and every row split."""
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


@pytest.fixture(scope="module")
def scene():
    """12x12 TX positions x 3 RX offsets (bistatic), 77 samples, 12x10x8 voxels."""
    cfg = synth.Config(1, 12, 2e-3, 77, (12, 10, 8), 2e-3, 20.0, "points", 5)
    sc = synth.make_scene(1, cfg=cfg, seed=31)
    rxo = np.array([[0.0, 0.0, 0.0], [-3.1e-3, 0.0, 0.0], [0.0, 2.6e-3, 0.0]])
    rng = np.random.default_rng(2)
    rows = np.repeat(sc.ant, 3, axis=0)
    meas = (rng.normal(size=(len(rows), cfg.n_freq)) +
            1j * rng.normal(size=(len(rows), cfg.n_freq))).astype(np.complex64)
    return sc, rows, meas, rxo


@pytest.mark.parametrize("K", [8, 16, 32, 64])
@pytest.mark.parametrize("is_ti", [False, True])
def test_ragged_bistatic_ti_vs_fp64(scene, oracle, K, is_ti):
    """F = 77 is no multiple of any K: the tail block runs every time."""
    sc, rows, meas, rxo = scene
    img, mom = ops.backproject(T(meas, np.complex64), T(rows, np.float32),
                               T(sc.x_locs, np.float32), T(sc.y_locs, np.float32),
                               T(sc.z_locs, np.float32), T(rxo, np.float32), f0=sc.f0,
                               df=sc.df, variant="rec", K=K, moment=True, is_ti=is_ti)
    want, wmom = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df,
                               is_ti=is_ti, moment=True)
    assert rel(img.cpu().numpy(), want) <= 1e-5
    assert rel(mom.cpu().numpy().reshape(-1, 3), wmom) <= 1e-5


@pytest.mark.parametrize("K", [128, 256])
@pytest.mark.parametrize("F", [77, 192, 256, 320, 333])
def test_long_reanchor_intervals_vs_fp64(oracle, K, F):
    """Re-anchor intervals of 128 and 256 samples (64-sample Horner chains
    combined in r^64): sample counts below one interval, exact multiples,
    and partial last intervals whose remainder is empty or not, bistatic
    with the TI offset."""
    cfg = synth.Config(1, 10, 2e-3, F, (10, 9, 7), 2e-3, 20.0, "points", 5)
    sc = synth.make_scene(1, cfg=cfg, seed=41)
    rxo = np.array([[0.0, 0.0, 0.0], [-3.1e-3, 0.0, 0.0], [0.0, 2.6e-3, 0.0]])
    rng = np.random.default_rng(F)
    rows = np.repeat(sc.ant, 3, axis=0)
    meas = (rng.normal(size=(len(rows), F)) + 1j * rng.normal(size=(len(rows), F))
            ).astype(np.complex64)
    img, mom = ops.backproject(T(meas, np.complex64), T(rows, np.float32),
                               T(sc.x_locs, np.float32), T(sc.y_locs, np.float32),
                               T(sc.z_locs, np.float32), T(rxo, np.float32), f0=sc.f0,
                               df=sc.df, variant="rec", K=K, moment=True, is_ti=True)
    want, wmom = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df,
                               is_ti=True, moment=True)
    assert rel(img.cpu().numpy(), want) <= 1e-5
    assert rel(mom.cpu().numpy().reshape(-1, 3), wmom) <= 1e-5


def test_single_and_repeated_voxels(scene, oracle):
    sc, rows, meas, rxo = scene
    args = (T(meas, np.complex64), T(rows, np.float32), T(sc.x_locs, np.float32),
            T(sc.y_locs, np.float32), T(sc.z_locs, np.float32), T(rxo, np.float32))
    for idx in ([417], [3, 3, 3, 959, 3]):
        vi = np.array(idx, np.int32)
        got = ops.backproject(*args, f0=sc.f0, df=sc.df, vox_idx=T(vi, np.int32)).cpu().numpy()
        want = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df,
                             vidx=vi)
        assert rel(got, want) <= 1e-5
        if len(idx) > 1:  # repeated voxels give identical values
            assert got[0] == got[1] == got[2] == got[4]
    empty = ops.backproject(*args, f0=sc.f0, df=sc.df, vox_idx=T(np.zeros(0, np.int32), np.int32))
    assert empty.numel() == 0


@pytest.mark.parametrize("rs", [1, 2, 7, 32])
def test_every_row_split_agrees(scene, oracle, rs):
    """Forced row-segment counts (the partial sums are added in FP64 in a
    fixed order) against the FP64 oracle."""
    sc, rows, meas, rxo = scene
    os.environ["MMN_REC_RS"] = str(rs)
    try:
        img = ops.backproject(T(meas, np.complex64), T(rows, np.float32),
                              T(sc.x_locs, np.float32), T(sc.y_locs, np.float32),
                              T(sc.z_locs, np.float32), T(rxo, np.float32), f0=sc.f0, df=sc.df)
    finally:
        del os.environ["MMN_REC_RS"]
    want = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df)
    assert rel(img.cpu().numpy(), want) <= 1e-5


@pytest.mark.parametrize("vpt,split,K", [(1, 1, 64), (1, 2, 16), (4, 1, 64), (2, 2, 32),
                                         (1, 1, 256), (4, 1, 128)])
def test_lane_shapes_agree(scene, oracle, vpt, split, K):
    """The recurrence's other lane shapes (voxels per lane, even/odd Horner
    chains; MMN_REC_VPT / MMN_REC_SPLIT) on the ragged bistatic TI inputs,
    with and without the moment, against the FP64 oracle."""
    sc, rows, meas, rxo = scene
    os.environ["MMN_REC_VPT"], os.environ["MMN_REC_SPLIT"] = str(vpt), str(split)
    try:
        img, mom = ops.backproject(T(meas, np.complex64), T(rows, np.float32),
                                   T(sc.x_locs, np.float32), T(sc.y_locs, np.float32),
                                   T(sc.z_locs, np.float32), T(rxo, np.float32), f0=sc.f0,
                                   df=sc.df, K=K, moment=True)
    finally:
        del os.environ["MMN_REC_VPT"], os.environ["MMN_REC_SPLIT"]
    want, wmom = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, rows, meas, rxo, sc.f0, sc.df,
                               moment=True)
    assert rel(img.cpu().numpy(), want) <= 1e-5
    assert rel(mom.cpu().numpy().reshape(-1, 3), wmom) <= 1e-5


def test_interp_bins_at_bin_boundaries(oracle):
    """Voxels placed where d / fft_spacing / 2 lands on or next to an integer,
    one antenna, a spectrum whose bins all differ: the MI355X interpolated
    kernel picks the reference's bin for every voxel (its fast bin falls back
    to the exact division there)."""
    fs = np.float32(3e8 / (2 * 3.0009e9) * 512 / 4096)
    S = 4096
    ks = np.arange(20, 90)
    base = (0.3 - ks * fs).astype(np.float32)
    z = np.concatenate([base, np.nextafter(base, np.float32(1)), np.nextafter(base, np.float32(-1))])
    z = np.sort(z.astype(np.float32))
    x = np.zeros(1, np.float32)
    ant = np.array([[0.0, 0.0, 0.3]], np.float32)
    spec = (np.arange(S) + 1.0 + 0.5j * np.arange(S)).astype(np.complex64)[None, :]
    lam = np.float32(3e8 / 77e9)
    args = (T(spec, np.complex64), T(ant, np.float32), T(x, np.float32), T(x, np.float32),
            T(z, np.float32))
    fast = ops.backproject_interp(*args, wavelength=lam, fft_spacing=fs).cpu().numpy().ravel()
    ref = ops.cuda_image(*args, None, slope=0.0, wavelength=lam, fft_spacing=fs,
                         interpolated=True).cpu().numpy().ravel()
    want = oracle.bp_ref(x, x, z, ant, spec, np.zeros((1, 3)), 0.0, lam, fs,
                         interp=True).ravel()
    assert np.allclose(np.abs(ref), np.abs(want), rtol=1e-6)
    # same bin <=> same magnitude (|m| differs between bins)
    assert np.allclose(np.abs(fast), np.abs(ref), rtol=1e-6)
    assert np.max(np.abs(fast - ref) / np.abs(ref)) <= 2e-6


def test_row_split_with_the_queried_workspace_vs_fp64(oracle):
    """128x128x80 voxels (2560 blocks of 512, fewer than 8 rounds), 64 rows:
    the launch splits every block into two row segments.  With exactly the
    workspace the size query gives, with one byte less (no split), and with
    none, the image and moment agree with O64 on a voxel sample, and the split
    and unsplit images with each other to rounding."""
    from mmnorm_amd import _lib
    cfg = synth.Config(1, 8, 2e-3, 16, (128, 128, 80), 1e-3, 20.0, "points", 5)
    sc = synth.make_scene(1, cfg=cfg, seed=7)
    n_grid = 128 * 128 * 80
    lib = _lib.load()
    ws_bytes = int(lib.mmn_bp_forward_workspace_size(n_grid, len(sc.ant), 1))
    assert ws_bytes == 2 * n_grid * 32
    meas, ant = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = T(sc.x_locs, np.float32), T(sc.y_locs, np.float32), T(sc.z_locs, np.float32)
    rx = torch.zeros((1, 3), dtype=torch.float32, device=dev())
    rng = np.random.default_rng(3)
    vi = np.unique(np.concatenate([[0, n_grid - 1], rng.choice(n_grid, 200, replace=False)]))
    want, wmom = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas,
                               np.zeros((1, 3)), sc.f0, sc.df, vidx=vi, moment=True)
    sel = torch.from_numpy(vi).to(dev())
    imgs = []
    for nbytes in (ws_bytes, ws_bytes - 1, 0):
        img = torch.empty(n_grid, dtype=torch.complex64, device=dev())
        mom = torch.empty((n_grid, 3), dtype=torch.complex64, device=dev())
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev())
        _lib.call("mmn_bp_forward", img.data_ptr(), mom.data_ptr(), xl.data_ptr(),
                  yl.data_ptr(), zl.data_ptr(), None, 0, ant.data_ptr(), meas.data_ptr(),
                  rx.data_ptr(), float(sc.f0), float(sc.df), 128, 128, 80, len(sc.ant), 1,
                  sc.meas.shape[1], 0, 2, 64, ws.data_ptr() if nbytes else None, nbytes,
                  _lib.stream_handle(dev()))
        torch.cuda.synchronize()
        assert rel(img[sel].cpu().numpy(), want) <= 1e-5
        assert rel(mom[sel].cpu().numpy(), wmom) <= 1e-5
        imgs.append(img.cpu().numpy())
    assert rel(imgs[0], imgs[2]) <= 1e-6
    # too small a workspace and none both run unsplit: the same bits
    assert np.array_equal(imgs[1], imgs[2])


def test_interp_sqrt_equals_sqrtf_on_every_float():
    """The interpolated branch's distances use a shorter correctly rounded
    square root (hardware root and a neighbour test); over all 2^32 float bit
    patterns it returns what sqrtf returns, so the distances, and with them the
    range bins, stay the reference's (imaging_gpu.cu:83-88)."""
    from mmnorm_amd import _lib
    res = torch.tensor([0, -1], dtype=torch.int64, device=dev())
    _lib.call("mmn_sqrt_rn_check", 0, 1 << 32, res.data_ptr(), _lib.stream_handle(dev()))
    bad, first = res.cpu().tolist()
    assert bad == 0, f"{bad} patterns differ from sqrtf, the first 0x{first & 0xffffffff:08x}"


def test_sqrt_check_rejects_a_range_past_the_patterns():
    from mmnorm_amd import _lib
    res = torch.tensor([0, -1], dtype=torch.int64, device=dev())
    with pytest.raises(Exception, match="2\\^32"):
        _lib.call("mmn_sqrt_rn_check", 1, 1 << 32, res.data_ptr(), _lib.stream_handle(dev()))
