"""Backward kernel: adjoint identity, FP64 central finite differences, and
torch.autograd on a float64 PyTorch reference of the same forward.

Loss (config 4): L = sum_v w_v |I_v|^2, w_v ~ U(0,1), I = scale * sum.
"""
import math

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


def small_case(seed=1, n_ap=8, F=32, grid=(8, 8, 6)):
    cfg = synth.Config(4, n_ap, 2e-3, F, grid, 2e-3, 20.0, "ellipsoid_box", 300)
    sc = synth.make_scene(4, cfg=cfg, seed=seed)
    return sc


def test_adjoint_identity():
    sc = small_case(F=64, grid=(16, 12, 10))
    rng = np.random.default_rng(4)
    x = T(rng.normal(size=sc.meas.shape) + 1j * rng.normal(size=sc.meas.shape), np.complex64)
    a = T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    n = len(sc.x_locs) * len(sc.y_locs) * len(sc.z_locs)
    g = T(rng.normal(size=n) + 1j * rng.normal(size=n), np.complex64)
    Ax = ops.backproject(x, a, xl, yl, zl, f0=sc.f0, df=sc.df).reshape(-1)
    AHg, _ = ops.backproject_backward(g, x, a, xl, yl, zl, f0=sc.f0, df=sc.df, need_ant=False)
    lhs = torch.vdot(g.to(torch.complex128), Ax.to(torch.complex128)).item()
    rhs = torch.vdot(AHg.to(torch.complex128).reshape(-1),
                     x.to(torch.complex128).reshape(-1)).item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_position_gradient_vs_central_fd(oracle):
    """dL/dp against FP64 central differences, h = 1e-6 m, on 64 coordinates."""
    sc = small_case()
    n = len(sc.x_locs) * len(sc.y_locs) * len(sc.z_locs)
    rng = np.random.default_rng(7)
    w = rng.uniform(size=n)
    scale = 1.0 / len(sc.ant)
    # the FD perturbs FP64 positions that are exactly representable in FP32
    ant = sc.ant.astype(np.float32).astype(np.float64)
    meas = T(sc.meas, np.complex64)
    a = T(ant, np.float32).requires_grad_(True)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    img = ops.backproject_ad(meas, a, xl, yl, zl, f0=sc.f0, df=sc.df, scale=scale)
    loss = (T(w, np.float32).reshape(img.shape) * img.abs() ** 2).sum()
    loss.backward()
    grad = a.grad.cpu().numpy()
    h = 1e-6
    coords = [(int(rng.integers(len(ant))), int(rng.integers(3))) for _ in range(64)]
    fd = []
    for (i, k) in coords:
        # the oracle reads FP32 positions, so the shifted row is recomputed in FP64
        fd.append((_loss64_shift(oracle, sc, ant, i, k, h, w, scale)
                   - _loss64_shift(oracle, sc, ant, i, k, -h, w, scale)) / (2 * h))
    fd = np.array(fd)
    got = np.array([grad[i, k] for (i, k) in coords])
    assert np.max(np.abs(got - fd)) <= 2e-3 * np.max(np.abs(fd))


def _loss64_shift(oracle, sc, ant, i, k, h, w, scale):
    """FP64 loss with antenna i moved by h along axis k (exact in FP64)."""
    X, Y, Z = np.meshgrid(sc.x_locs.astype(np.float32).astype(np.float64),
                          sc.y_locs.astype(np.float32).astype(np.float64),
                          sc.z_locs.astype(np.float32).astype(np.float64), indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    base = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, ant, sc.meas, np.zeros((1, 3)),
                         sc.f0, sc.df)
    f = sc.f0 + sc.df * np.arange(sc.meas.shape[1])
    s = sc.meas[i].astype(np.complex128)

    def row(p):
        d = 2 * np.linalg.norm(v - p[None, :], axis=1)
        return np.exp(-2j * np.pi * d[:, None] * f[None, :] / C0) @ s

    p0 = ant[i].copy()
    p1 = p0.copy()
    p1[k] += h
    img = base - row(p0) + row(p1)
    return float(np.sum(w * np.abs(scale * img) ** 2))


def _row_block64(meas, ant, v, f0, df, scale):
    """Pure PyTorch float64 forward of a block of rows:
    scale * sum_ij s_ij exp(-j 2 pi f_j d_iv / c) over the block."""
    F = meas.shape[1]
    k = 2 * math.pi * (f0 + df * torch.arange(F, dtype=torch.float64, device=meas.device)) / C0
    d = 2 * torch.linalg.norm(v[:, None, :] - ant[None, :, :], dim=2)          # (V, a)
    ph = torch.exp(-1j * (d[..., None] * k).to(torch.complex128))                  # (V, a, F)
    return scale * (ph * meas[None]).sum(dim=(1, 2))


def _torch_reference_grads(meas, ant, v, w, f0, df, scale, chunk=25):
    """torch.autograd gradients of L = sum w |I|^2, row block by row block.

    L depends on a block's rows only through I, so each block's gradient is
    that of Re(<G, I_block>) with G = dL/dI = 2 w I held fixed.
    """
    with torch.no_grad():
        I = sum(_row_block64(meas[lo:lo + chunk], ant[lo:lo + chunk], v, f0, df, scale)
                for lo in range(0, ant.shape[0], chunk))
    G = 2 * w * I
    gm = torch.zeros_like(meas)
    ga = torch.zeros_like(ant)
    for lo in range(0, ant.shape[0], chunk):
        m = meas[lo:lo + chunk].detach().clone().requires_grad_(True)
        a = ant[lo:lo + chunk].detach().clone().requires_grad_(True)
        blk = _row_block64(m, a, v, f0, df, scale)
        torch.real(torch.vdot(G, blk)).backward()
        gm[lo:lo + chunk] = m.grad
        ga[lo:lo + chunk] = a.grad
    return gm, ga


def test_gradients_vs_torch_autograd_reference():
    """Config-4 aperture and band (40x40, 128 samples) on a 16^3 sub-grid."""
    sc = synth.make_scene(4, device=dev())
    sub = (slice(40, 56), slice(40, 56), slice(40, 56))
    xs, ys, zs = sc.x_locs[sub[0]], sc.y_locs[sub[1]], sc.z_locs[sub[2]]
    n = 16 ** 3
    rng = np.random.default_rng(3)
    w = rng.uniform(size=n)
    scale = 1.0 / len(sc.ant)
    meas = T(sc.meas, np.complex64).requires_grad_(True)
    a = T(sc.ant, np.float32).requires_grad_(True)
    xl, yl, zl = (T(v, np.float32) for v in (xs, ys, zs))
    img = ops.backproject_ad(meas, a, xl, yl, zl, f0=sc.f0, df=sc.df, scale=scale)
    wt = T(w, np.float32).reshape(img.shape)
    (wt * img.abs() ** 2).sum().backward()

    X, Y, Z = torch.meshgrid(xl.double(), yl.double(), zl.double(), indexing="ij")
    v = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1)
    gm64, ga64 = _torch_reference_grads(meas.detach().to(torch.complex128),
                                        a.detach().double(), v,
                                        torch.from_numpy(w).to(dev()), sc.f0, sc.df, scale)
    gm, ga = meas.grad.to(torch.complex128), a.grad.double()
    assert (torch.linalg.norm(gm - gm64) / torch.linalg.norm(gm64)).item() < 1e-4
    assert (torch.linalg.norm(ga - ga64) / torch.linalg.norm(ga64)).item() < 1e-3


def test_backward_deterministic():
    sc = small_case(F=64, grid=(12, 12, 8))
    rng = np.random.default_rng(8)
    meas, a = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    n = 12 * 12 * 8
    g = T(rng.normal(size=n) + 1j * rng.normal(size=n), np.complex64)
    r1 = ops.backproject_backward(g, meas, a, xl, yl, zl, f0=sc.f0, df=sc.df)
    r2 = ops.backproject_backward(g, meas, a, xl, yl, zl, f0=sc.f0, df=sc.df)
    assert torch.equal(r1[0], r2[0]) and torch.equal(r1[1], r2[1])
