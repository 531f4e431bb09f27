"""GPU parity of the backprojection and vote kernels against the CPU oracle.

Tolerances (relative L2 over the compared voxels unless stated):
* reference-arithmetic mirror vs the C restatement of imaging_gpu.cu: same
  operation order, no contraction, FP64 cos/sin from different libraries
  (OCML vs glibc) narrowed to FP32 -> identical but for rare last-bit flips:
  max |diff| <= 1e-6 * max|I|.
* recurrence vs the FP64 oracle: <= 1e-5 (target of the path is 1e-4).
* plain FP32 sincosf vs the FP64 oracle: <= 1e-3 (FP32 phase of ~1000 rad).
"""
import numpy as np
import pytest
import torch

from mmnorm_amd import ops, synth

pytestmark = pytest.mark.gpu

C0 = synth.C0


def dev():
    return torch.device("cuda", 0)


def T(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev())


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def sc1():
    return synth.make_scene(1)


@pytest.fixture(scope="module")
def sub1(sc1):
    """Config-1 scene on a 16 x 16 x 8 sub-grid around the scene."""
    xl, yl, zl = sc1.x_locs[::4], sc1.y_locs[::4], sc1.z_locs[::4]
    return sc1, xl, yl, zl


def _dev_inputs(sc, xl, yl, zl, rxo=None):
    rxo = np.zeros((1, 3)) if rxo is None else rxo
    return (T(sc.meas, np.complex64), T(sc.ant, np.float32), T(xl, np.float32),
            T(yl, np.float32), T(zl, np.float32), T(rxo, np.float32))


def test_mirror_full_matches_reference_restatement(sub1, oracle):
    sc, xl, yl, zl = sub1
    m, a, x, y, z, r = _dev_inputs(sc, xl, yl, zl)
    got = ops.cuda_image(m, a, x, y, z, r, slope=np.float32(sc.slope),
                         wavelength=np.float32(sc.wavelength)).cpu().numpy().ravel()
    want = oracle.bp_ref(xl, yl, zl, sc.ant, sc.meas, np.zeros((1, 3)), sc.slope,
                         sc.wavelength)
    assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want))
    same = np.mean((got.real == want.real) & (got.imag == want.imag))
    assert same > 0.9, f"only {same:.3f} of voxels bit-identical"


@pytest.mark.parametrize("K", [8, 16, 32, 64])
def test_rec_matches_fp64_oracle(sub1, oracle, K):
    sc, xl, yl, zl = sub1
    m, a, x, y, z, r = _dev_inputs(sc, xl, yl, zl)
    img, mom = ops.backproject(m, a, x, y, z, r, f0=sc.f0, df=sc.df, variant="rec", K=K,
                               moment=True)
    want, wmom = oracle.bp_f64(xl, yl, zl, sc.ant, sc.meas, np.zeros((1, 3)), sc.f0, sc.df,
                               moment=True)
    assert rel(img.cpu().numpy(), want) <= 1e-5
    assert rel(mom.cpu().numpy().reshape(-1, 3), wmom) <= 1e-5


def test_plain_matches_fp64_oracle(sub1, oracle):
    sc, xl, yl, zl = sub1
    m, a, x, y, z, r = _dev_inputs(sc, xl, yl, zl)
    img, mom = ops.backproject(m, a, x, y, z, r, f0=sc.f0, df=sc.df, variant="plain",
                               moment=True)
    want, wmom = oracle.bp_f64(xl, yl, zl, sc.ant, sc.meas, np.zeros((1, 3)), sc.f0, sc.df,
                               moment=True)
    assert rel(img.cpu().numpy(), want) <= 1e-3
    assert rel(mom.cpu().numpy().reshape(-1, 3), wmom) <= 1e-3


def test_rec_voxel_list_equals_grid(sub1):
    sc, xl, yl, zl = sub1
    m, a, x, y, z, r = _dev_inputs(sc, xl, yl, zl)
    full, fm = ops.backproject(m, a, x, y, z, r, f0=sc.f0, df=sc.df, moment=True)
    rng = np.random.default_rng(5)
    idx = rng.choice(full.numel(), size=333, replace=False).astype(np.int32)
    part, pm = ops.backproject(m, a, x, y, z, r, f0=sc.f0, df=sc.df, moment=True,
                               vox_idx=T(idx, np.int32))
    assert torch.equal(part, full.reshape(-1)[torch.from_numpy(idx).long().to(dev())])
    assert torch.equal(pm, fm.reshape(-1, 3)[torch.from_numpy(idx).long().to(dev())])


@pytest.mark.parametrize("variant", ["rec", "plain"])
def test_point_scatterer_focus_closed_form(variant):
    """Noise-free scatterer at a voxel centre: every phasor cancels, I = rho * A * F."""
    cfg = synth.CONFIGS[1]
    ant = synth.aperture(12, 2e-3, 0.3)
    xl, yl, zl = synth.axis(8, 2e-3), synth.axis(8, 2e-3), synth.axis(4, 2e-3)
    pos = np.array([[xl[3], yl[5], zl[1]]], dtype=np.float32).astype(np.float64)
    rho = np.array([0.7 - 0.4j])
    freqs = cfg.f0 + cfg.df * np.arange(cfg.n_freq)
    # distances from the FP32-rounded geometry the kernel sees
    s = synth.synthesize(ant.astype(np.float32).astype(np.float64), pos, rho, freqs)
    m, a = T(s, np.complex64), T(ant, np.float32)
    img = ops.backproject(m, a, T(xl, np.float32), T(yl, np.float32), T(zl, np.float32),
                          f0=cfg.f0, df=cfg.df, variant=variant).cpu().numpy()
    peak = img[3, 5, 1]
    want = rho[0] * len(ant) * cfg.n_freq
    tol = 2e-5 if variant == "rec" else 2e-3
    assert abs(peak - want) <= tol * abs(want)
    assert np.unravel_index(np.argmax(np.abs(img)), img.shape) == (3, 5, 1)


def test_single_row_single_sample_magnitude():
    """One row, one sample: |I_v| = |s| for every voxel, phase = -2 pi f0 d / c."""
    xl, yl, zl = synth.axis(4, 3e-3), synth.axis(3, 3e-3), synth.axis(2, 3e-3)
    ant = np.array([[0.01, -0.02, 0.3]])
    s = np.array([[1.5 + 2.0j]])
    f0 = 77e9
    img = ops.backproject(T(s, np.complex64), T(ant, np.float32), T(xl, np.float32),
                          T(yl, np.float32), T(zl, np.float32), f0=f0, df=1e6,
                          variant="rec").cpu().numpy()
    assert np.allclose(np.abs(img), 2.5, rtol=1e-6)
    X, Y, Z = np.meshgrid(xl.astype(np.float32), yl.astype(np.float32),
                          zl.astype(np.float32), indexing="ij")
    a32 = ant.astype(np.float32).astype(np.float64)[0]
    d = 2 * np.sqrt((X - a32[0]) ** 2 + (Y - a32[1]) ** 2 + (Z - a32[2]) ** 2)
    want = s[0, 0] * np.exp(-2j * np.pi * f0 * d / C0)
    assert np.max(np.abs(img - want)) <= 2e-5 * 2.5


def test_interp_and_vote_mirror_match_restatement(oracle):
    """Range-bin branch and the reference vote on 4096-bin spectra (4 RX, TI offset)."""
    rng = np.random.default_rng(11)
    n_tx, R, S = 24, 4, 4096
    lam = np.float32(C0 / 77.5001e9)
    rxo = np.array([[-0.005 - lam / 2 * k, 0, 0] for k in (3, 2, 1, 0)], np.float32)
    ant_tx = np.stack([rng.uniform(-0.05, 0.05, n_tx), rng.uniform(-0.05, 0.05, n_tx),
                       np.full(n_tx, 0.35)], 1)
    ant = np.repeat(ant_tx, R, axis=0)
    spec = (rng.normal(size=(n_tx * R, S)) + 1j * rng.normal(size=(n_tx * R, S)))
    xl, yl, zl = synth.axis(10, 2e-3), synth.axis(9, 2e-3), synth.axis(7, 2e-3)
    fs = np.float32(3e8 / (2 * 3.0009e9) * 512 / S)
    m, a = T(spec, np.complex64), T(ant, np.float32)
    x, y, z, r = T(xl, np.float32), T(yl, np.float32), T(zl, np.float32), T(rxo, np.float32)
    got = ops.cuda_image(m, a, x, y, z, r, slope=5861160.0, wavelength=lam, fft_spacing=fs,
                         is_ti=True, interpolated=True)
    want = oracle.bp_ref(xl, yl, zl, ant, spec, rxo, 5861160.0, lam, fs, is_ti=True,
                         interp=True)
    g = got.cpu().numpy().ravel()
    assert np.max(np.abs(g - want)) <= 1e-6 * np.max(np.abs(want))
    vidx = np.flatnonzero(np.abs(want) > 0.3 * np.abs(want).max()).astype(np.int32)
    sn = ops.sn_est(got.reshape(-1), T(vidx, np.int32), m, a, x, y, z, r, wavelength=lam,
                    fft_spacing=fs, is_ti=True).cpu().numpy()
    wsn = oracle.sn_ref(vidx, g, xl, yl, zl, ant, spec, rxo, lam, fs, is_ti=True)
    assert np.max(np.abs(sn - wsn)) <= 1e-5 * np.max(np.abs(wsn))


def test_vote_from_moment_matches_fp64(sub1, oracle):
    sc, xl, yl, zl = sub1
    m, a, x, y, z, r = _dev_inputs(sc, xl, yl, zl)
    img = ops.backproject(m, a, x, y, z, r, f0=sc.f0, df=sc.df)
    mag = img.abs().reshape(-1)
    vidx = torch.nonzero(mag > 0.05 * mag.max()).reshape(-1).to(torch.int32)
    _, mom = ops.backproject(m, a, x, y, z, r, f0=sc.f0, df=sc.df, moment=True, vox_idx=vidx)
    n = ops.sn_from_moment(img.reshape(-1), mom, vidx).cpu().numpy()
    wi, wm = oracle.bp_f64(xl, yl, zl, sc.ant, sc.meas, np.zeros((1, 3)), sc.f0, sc.df,
                           vidx=vidx.cpu().numpy(), moment=True)
    wn = (wi.real[:, None] * wm.real + wi.imag[:, None] * wm.imag) / np.abs(wi)[:, None]
    cosang = np.sum(n * wn, 1) / np.linalg.norm(n, axis=1) / np.linalg.norm(wn, axis=1)
    assert np.degrees(np.arccos(np.clip(cosang.min(), -1, 1))) < 0.05


def test_single_row_vote_points_at_antenna():
    """One row: the normal is the unit vector from the voxel to the (virtual) antenna."""
    xl, yl, zl = synth.axis(5, 2e-3), synth.axis(5, 2e-3), synth.axis(3, 2e-3)
    ant = np.array([[0.004, -0.006, 0.3]])
    rng = np.random.default_rng(3)
    s = rng.normal(size=(1, 16)) + 1j * rng.normal(size=(1, 16))
    x, y, z = T(xl, np.float32), T(yl, np.float32), T(zl, np.float32)
    m, a = T(s, np.complex64), T(ant, np.float32)
    img = ops.backproject(m, a, x, y, z, f0=77e9, df=4e9 / 15)
    vidx = torch.arange(img.numel(), dtype=torch.int32, device=dev())
    _, mom = ops.backproject(m, a, x, y, z, f0=77e9, df=4e9 / 15, moment=True, vox_idx=vidx)
    n = ops.sn_from_moment(img.reshape(-1), mom, vidx).cpu().numpy()
    X, Y, Z = np.meshgrid(xl, yl, zl, indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    want = ant[0] - v
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    got = n / np.linalg.norm(n, axis=1, keepdims=True)
    assert np.max(np.abs(got - want)) < 1e-5


def test_symmetric_aperture_normal_is_vertical():
    """Square aperture centred over a point at the origin: the focus normal is (0,0,1)."""
    ant = synth.aperture(16, 2e-3, 0.3)
    pos = np.zeros((1, 3))
    freqs = 77e9 + (4e9 / 31) * np.arange(32)
    s = synth.synthesize(ant, pos, np.array([1.0 + 0j]), freqs)
    xl = synth.axis(3, 1e-3)
    x = T(xl, np.float32)
    m, a = T(s, np.complex64), T(ant, np.float32)
    img = ops.backproject(m, a, x, x, x, f0=77e9, df=4e9 / 31)
    vidx = torch.tensor([13], dtype=torch.int32, device=dev())  # (1,1,1) = origin
    _, mom = ops.backproject(m, a, x, x, x, f0=77e9, df=4e9 / 31, moment=True, vox_idx=vidx)
    n = ops.sn_from_moment(img.reshape(-1), mom, vidx).cpu().numpy()[0]
    n /= np.linalg.norm(n)
    assert abs(n[0]) < 1e-5 and abs(n[1]) < 1e-5 and n[2] > 0.99999


def test_linearity_and_row_permutation(sub1):
    sc, xl, yl, zl = sub1
    m, a, x, y, z, r = _dev_inputs(sc, xl, yl, zl)
    rng = np.random.default_rng(2)
    m2 = T(rng.normal(size=sc.meas.shape) + 1j * rng.normal(size=sc.meas.shape), np.complex64)
    i1 = ops.backproject(m, a, x, y, z, f0=sc.f0, df=sc.df)
    i2 = ops.backproject(m2, a, x, y, z, f0=sc.f0, df=sc.df)
    i12 = ops.backproject((2 * m - 3j * m2).contiguous(), a, x, y, z, f0=sc.f0, df=sc.df)
    assert rel((i12).cpu().numpy(), (2 * i1 - 3j * i2).cpu().numpy()) < 1e-5
    perm = torch.from_numpy(rng.permutation(m.shape[0])).to(dev())
    ip = ops.backproject(m[perm].contiguous(), a[perm].contiguous(), x, y, z, f0=sc.f0,
                         df=sc.df)
    assert rel(ip.cpu().numpy(), i1.cpu().numpy()) < 1e-5


def test_bad_arguments_raise():
    x = torch.zeros(4, dtype=torch.float32, device=dev())
    m = torch.zeros((3, 8), dtype=torch.complex64, device=dev())
    a = torch.zeros((2, 3), dtype=torch.float32, device=dev())
    with pytest.raises(ValueError):
        ops.backproject(m, a, x, x, x, f0=77e9, df=1e6)
    a3 = torch.zeros((3, 3), dtype=torch.float32, device=dev())
    with pytest.raises(ValueError):
        ops.backproject(m, a3, x, x, x, f0=77e9, df=1e6, variant="rec", K=12)
    with pytest.raises(TypeError):
        ops.backproject(m.to(torch.complex128), a3, x, x, x, f0=77e9, df=1e6)
