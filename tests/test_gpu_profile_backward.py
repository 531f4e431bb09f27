"""The range-profile backward (mmn_bp_profile_backward): the adjoint identity
with the profile forward, FP64 central differences, torch.autograd on a
float64 PyTorch reference, the FP64 NumPy backward on ragged bistatic TI
inputs and voxel lists, single outputs, chunk splits, determinism, and the
config-4 gradients against the recurrence backward."""
import os

import numpy as np
import pytest
import torch

from mmnorm_amd import ops, synth
from test_gpu_backward import (_backward64, _loss64_shift, _torch_reference_grads,
                               small_case)

pytestmark = pytest.mark.gpu


def dev():
    return torch.device("cuda", 0)


def T(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev())


def rel(a, b):
    return float(torch.linalg.norm((a - b).reshape(-1)) / torch.linalg.norm(b.reshape(-1)))


def test_adjoint_identity_with_profile_forward():
    sc = small_case(F=64, grid=(16, 12, 10))
    rng = np.random.default_rng(4)
    x = T(rng.normal(size=sc.meas.shape) + 1j * rng.normal(size=sc.meas.shape), np.complex64)
    a = T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    n = len(sc.x_locs) * len(sc.y_locs) * len(sc.z_locs)
    g = T(rng.normal(size=n) + 1j * rng.normal(size=n), np.complex64)
    Ax = ops.backproject(x, a, xl, yl, zl, f0=sc.f0, df=sc.df, variant="profile").reshape(-1)
    AHg, _ = ops.backproject_backward(g, x, a, xl, yl, zl, f0=sc.f0, df=sc.df, need_ant=False,
                                      variant="profile")
    lhs = torch.vdot(g.to(torch.complex128), Ax.to(torch.complex128)).item()
    rhs = torch.vdot(AHg.to(torch.complex128).reshape(-1),
                     x.to(torch.complex128).reshape(-1)).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_position_gradient_vs_central_fd(oracle):
    """dL/dp of the profile pair against FP64 central differences of the
    exact sum, h = 1e-6 m, 64 coordinates."""
    sc = small_case()
    n = len(sc.x_locs) * len(sc.y_locs) * len(sc.z_locs)
    rng = np.random.default_rng(7)
    w = rng.uniform(size=n)
    scale = 1.0 / len(sc.ant)
    ant = sc.ant.astype(np.float32).astype(np.float64)
    meas = T(sc.meas, np.complex64)
    a = T(ant, np.float32).requires_grad_(True)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    img = ops.backproject_ad(meas, a, xl, yl, zl, f0=sc.f0, df=sc.df, scale=scale,
                             variant="profile")
    (T(w, np.float32).reshape(img.shape) * img.abs() ** 2).sum().backward()
    grad = a.grad.cpu().numpy()
    h = 1e-6
    coords = [(int(rng.integers(len(ant))), int(rng.integers(3))) for _ in range(64)]
    fd = np.array([(_loss64_shift(oracle, sc, ant, i, k, h, w, scale)
                    - _loss64_shift(oracle, sc, ant, i, k, -h, w, scale)) / (2 * h)
                   for (i, k) in coords])
    got = np.array([grad[i, k] for (i, k) in coords])
    assert np.max(np.abs(got - fd)) <= 2e-3 * np.max(np.abs(fd))


def test_gradients_vs_torch_autograd_reference():
    """Config-4 aperture and band (40x40, 128 samples) on a 16^3 sub-grid."""
    sc = synth.make_scene(4, device=dev())
    sub = (slice(40, 56), slice(40, 56), slice(40, 56))
    xs, ys, zs = sc.x_locs[sub[0]], sc.y_locs[sub[1]], sc.z_locs[sub[2]]
    rng = np.random.default_rng(3)
    w = rng.uniform(size=16 ** 3)
    scale = 1.0 / len(sc.ant)
    meas = T(sc.meas, np.complex64).requires_grad_(True)
    a = T(sc.ant, np.float32).requires_grad_(True)
    xl, yl, zl = (T(v, np.float32) for v in (xs, ys, zs))
    img = ops.backproject_ad(meas, a, xl, yl, zl, f0=sc.f0, df=sc.df, scale=scale,
                             variant="profile")
    (T(w, np.float32).reshape(img.shape) * img.abs() ** 2).sum().backward()
    X, Y, Z = torch.meshgrid(xl.double(), yl.double(), zl.double(), indexing="ij")
    v = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1)
    gm64, ga64 = _torch_reference_grads(meas.detach().to(torch.complex128),
                                        a.detach().double(), v,
                                        torch.from_numpy(w).to(dev()), sc.f0, sc.df, scale)
    assert rel(meas.grad.to(torch.complex128), gm64) < 1e-4
    assert rel(a.grad.double(), ga64) < 1e-3


@pytest.mark.parametrize("n_list", [1, 77, 129, None])
def test_ragged_bistatic_ti_vs_fp64(n_list):
    sc = small_case(seed=5, n_ap=6, F=40, grid=(9, 7, 5))
    rng = np.random.default_rng(11)
    rxo = np.array([[0.0, 0.0, 0.0], [4e-3, -1e-3, 0.0]])
    meas = np.repeat(sc.meas, 2, axis=0)[: 2 * len(sc.ant)]
    meas = meas * np.exp(1j * rng.uniform(0, 2 * np.pi, size=meas.shape))
    ant = np.repeat(sc.ant, 2, axis=0).astype(np.float32).astype(np.float64)
    X, Y, Z = np.meshgrid(sc.x_locs.astype(np.float32).astype(np.float64),
                          sc.y_locs.astype(np.float32).astype(np.float64),
                          sc.z_locs.astype(np.float32).astype(np.float64), indexing="ij")
    vall = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    if n_list is None:
        idx, v = None, vall
    else:
        sel = np.sort(rng.choice(len(vall), size=n_list, replace=False)).astype(np.int32)
        idx, v = T(sel, np.int32), vall[sel]
    g = rng.normal(size=len(v)) + 1j * rng.normal(size=len(v))
    scale = 1.0 / len(ant)
    xl, yl, zl = (T(q, np.float32) for q in (sc.x_locs, sc.y_locs, sc.z_locs))
    gm, ga = ops.backproject_backward(T(g, np.complex64), T(meas, np.complex64),
                                      T(ant, np.float32), xl, yl, zl, T(rxo, np.float32),
                                      f0=sc.f0, df=sc.df, scale=scale, vox_idx=idx,
                                      is_ti=True, variant="profile")
    gm64, ga64 = _backward64(g, meas.astype(np.complex64).astype(np.complex128), ant,
                             rxo.astype(np.float32).astype(np.float64), v, sc.f0, sc.df,
                             scale, True)
    gm, ga = gm.cpu().numpy().astype(np.complex128), ga.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(gm - gm64) <= 1e-5 * np.linalg.norm(gm64)
    assert np.linalg.norm(ga - ga64) <= 1e-4 * np.linalg.norm(ga64)


@pytest.fixture(scope="module")
def mid_case():
    sc = small_case(seed=9, n_ap=10, F=64, grid=(20, 16, 12))
    rng = np.random.default_rng(12)
    n = 20 * 16 * 12
    g = T(rng.normal(size=n) + 1j * rng.normal(size=n), np.complex64)
    args = (T(sc.meas, np.complex64), T(sc.ant, np.float32),
            *(T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs)))
    return sc, g, args


def test_single_outputs_and_determinism(mid_case):
    sc, g, args = mid_case
    kw = dict(f0=sc.f0, df=sc.df, scale=0.5, variant="profile")
    gm, ga = ops.backproject_backward(g, *args, **kw)
    gm2, ga2 = ops.backproject_backward(g, *args, **kw)
    assert torch.equal(gm, gm2) and torch.equal(ga, ga2)
    only_m, none_a = ops.backproject_backward(g, *args, need_ant=False, **kw)
    none_m, only_a = ops.backproject_backward(g, *args, need_meas=False, **kw)
    assert none_a is None and none_m is None
    assert torch.equal(only_m, gm) and torch.equal(only_a, ga)


@pytest.mark.parametrize("chunks", ["1", "3", "64"])
def test_every_chunk_split(mid_case, chunks):
    sc, g, args = mid_case
    kw = dict(f0=sc.f0, df=sc.df, scale=0.5, variant="profile")
    gm, ga = ops.backproject_backward(g, *args, **kw)
    os.environ["MMN_PROFB_CHUNKS"] = chunks
    try:
        gm2, ga2 = ops.backproject_backward(g, *args, **kw)
    finally:
        del os.environ["MMN_PROFB_CHUNKS"]
    assert rel(gm2, gm) <= 1e-6 and rel(ga2, ga) <= 1e-5


def test_config4_matches_recurrence_backward():
    """Config 4 at full size (40x40 aperture, 128 samples, 96^3 voxels),
    L = sum w |I|^2: the profile pair's gradients against the recurrence
    pair's."""
    sc = synth.make_scene(4, device=dev())
    args = (T(sc.meas, np.complex64), T(sc.ant, np.float32),
            *(T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs)))
    scale = 1.0 / len(sc.ant)
    img = ops.backproject(*args, f0=sc.f0, df=sc.df).reshape(-1) * scale
    w = torch.rand(img.numel(), device=dev(), generator=torch.Generator(dev()).manual_seed(2))
    g = (2.0 * w * img).to(torch.complex64)
    ref = ops.backproject_backward(g, *args, f0=sc.f0, df=sc.df, scale=scale)
    got = ops.backproject_backward(g, *args, f0=sc.f0, df=sc.df, scale=scale,
                                   variant="profile")
    assert rel(got[0], ref[0]) <= 1e-5
    assert rel(got[1], ref[1]) <= 1e-4


def test_shuffled_full_list_matches_grid(mid_case):
    """Every voxel as a list in random order (the list path sorts it by brick
    and groups each brick's entries) against the grid path."""
    sc, g, args = mid_case
    kw = dict(f0=sc.f0, df=sc.df, scale=0.5, variant="profile")
    gm, ga = ops.backproject_backward(g, *args, **kw)
    perm = np.random.default_rng(5).permutation(g.numel()).astype(np.int32)
    pt = T(perm, np.int32)
    gm2, ga2 = ops.backproject_backward(g[pt.long()].contiguous(), *args, vox_idx=pt, **kw)
    assert rel(gm2, gm) <= 1e-6 and rel(ga2, ga) <= 1e-5


def test_descending_sweep_vs_fp64():
    """df < 0 (frequencies falling from f0): both outputs against the FP64
    NumPy backward."""
    sc = small_case(seed=6, n_ap=6, F=40, grid=(9, 7, 5))
    rng = np.random.default_rng(13)
    df = -sc.df
    f0 = sc.f0 + (sc.meas.shape[1] - 1) * sc.df
    meas = sc.meas * np.exp(1j * rng.uniform(0, 2 * np.pi, size=sc.meas.shape))
    ant = sc.ant.astype(np.float32).astype(np.float64)
    X, Y, Z = np.meshgrid(sc.x_locs.astype(np.float32).astype(np.float64),
                          sc.y_locs.astype(np.float32).astype(np.float64),
                          sc.z_locs.astype(np.float32).astype(np.float64), indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    g = rng.normal(size=len(v)) + 1j * rng.normal(size=len(v))
    xl, yl, zl = (T(q, np.float32) for q in (sc.x_locs, sc.y_locs, sc.z_locs))
    gm, ga = ops.backproject_backward(T(g, np.complex64), T(meas, np.complex64),
                                      T(ant, np.float32), xl, yl, zl, f0=f0, df=df, scale=0.1,
                                      variant="profile")
    gm64, ga64 = _backward64(g, meas.astype(np.complex64).astype(np.complex128), ant,
                             np.zeros((1, 3)), v, f0, df, 0.1, False)
    gm, ga = gm.cpu().numpy().astype(np.complex128), ga.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(gm - gm64) <= 1e-5 * np.linalg.norm(gm64)
    assert np.linalg.norm(ga - ga64) <= 1e-4 * np.linalg.norm(ga64)
