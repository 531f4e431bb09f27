"""The reference's real-data grid size, 550 x 600 x 250 voxels at 1 mm
(82.5 M voxels, the grid of image_processor.py:330-355), with a few rows:
every kernel that maps a whole grid -- the recurrence, the range profile, the
interpolated branch and the reference-arithmetic mirror -- against the
restatements on sampled voxels, the grid's corners and its last index among
them, so that the index arithmetic is exercised at full size."""
import numpy as np
import pytest
import torch

from mmnorm_amd import ops

pytestmark = pytest.mark.gpu

NX, NY, NZ = 550, 600, 250
C0 = 299792458.0


def dev():
    return torch.device("cuda", 0)


def T(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev())


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def grid():
    xl = ((np.arange(NX) - (NX - 1) / 2) * 1e-3).astype(np.float32)
    yl = ((np.arange(NY) - (NY - 1) / 2) * 1e-3).astype(np.float32)
    zl = (np.arange(NZ) * 1e-3).astype(np.float32)
    rng = np.random.default_rng(11)
    # 4 TX positions 0.5 m above the grid, each repeated for 4 RX offsets
    tx = np.array([[-0.01, -0.01, 0.5], [0.01, -0.01, 0.5], [-0.01, 0.01, 0.5],
                   [0.01, 0.01, 0.5]], np.float32)
    rxo = np.array([[0.0, 0.0, 0.0], [2.5e-3, 0.0, 0.0], [0.0, 2.5e-3, 0.0],
                    [2.5e-3, 2.5e-3, 0.0]], np.float32)
    rows = np.repeat(tx, 4, axis=0)
    V = NX * NY * NZ
    corners = [((x * NY) + y) * NZ + z for x in (0, NX - 1) for y in (0, NY - 1)
               for z in (0, NZ - 1)]
    sample = np.unique(np.concatenate([rng.integers(0, V, 3000), corners,
                                       [V - 1, V // 2]])).astype(np.int32)
    return xl, yl, zl, rows, rxo, rng, sample


def test_full_branch_recurrence_and_profile(grid, oracle):
    xl, yl, zl, rows, rxo, rng, sample = grid
    F, f0, df = 16, 77e9, 4e9 / 15
    meas = (rng.normal(size=(len(rows), F)) + 1j * rng.normal(size=(len(rows), F)))
    args = (T(meas, np.complex64), T(rows, np.float32), T(xl, np.float32), T(yl, np.float32),
            T(zl, np.float32), T(rxo, np.float32))
    idx = torch.from_numpy(sample.astype(np.int64)).to(dev())
    want = oracle.bp_f64(xl, yl, zl, rows, meas.astype(np.complex64), rxo, f0, df,
                         vidx=sample)
    for variant in ("rec", "profile"):
        img = ops.backproject(*args, f0=f0, df=df, variant=variant)
        assert img.shape == (NX, NY, NZ)
        got = img.reshape(-1)[idx].cpu().numpy()
        del img
        assert rel(got, want) <= 1e-5, variant


def test_interpolated_branch(grid, oracle):
    xl, yl, zl, rows, rxo, rng, sample = grid
    S = 4096
    fs = float(np.float32(3e8 / (2 * 3.0009e9) * 512 / S))
    lam = float(np.float32(3e8 / 77e9))
    spec = (rng.normal(size=(len(rows), S)) + 1j * rng.normal(size=(len(rows), S)))
    args = (T(spec, np.complex64), T(rows, np.float32), T(xl, np.float32), T(yl, np.float32),
            T(zl, np.float32), T(rxo, np.float32))
    idx = torch.from_numpy(sample.astype(np.int64)).to(dev())
    want = oracle.bp_ref(xl, yl, zl, rows, spec.astype(np.complex64), rxo, 0.0, lam, fs,
                         interp=True, vidx=sample)
    fast = ops.backproject_interp(*args, wavelength=lam, fft_spacing=fs)
    got = fast.reshape(-1)[idx].cpu().numpy()
    del fast
    assert rel(got, want) <= 1e-5
    mirror = ops.cuda_image(*args, slope=0.0, wavelength=lam, fft_spacing=fs, interpolated=True)
    got = mirror.reshape(-1)[idx].cpu().numpy()
    del mirror
    assert rel(got, want) <= 1e-6
