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


def _backward64(g, meas, ant, rxo, v, f0, df, scale, is_ti):
    """FP64 NumPy backward of I = scale * sum_ij s_ij exp(-j k_j d_iv), bistatic:
    d_iv = |v - p_i| + |v - p_i - o_i| (+ 0.15 m), k_j = 2 pi (f0 + j df) / c."""
    A, F = meas.shape
    k = 2 * np.pi * (f0 + df * np.arange(F)) / C0
    o = rxo[np.arange(A) % len(rxo)]
    e1 = v[None, :, :] - ant[:, None, :]                     # (A, V, 3)
    e2 = e1 - o[:, None, :]
    n1 = np.linalg.norm(e1, axis=2)
    n2 = np.linalg.norm(e2, axis=2)
    d = n1 + n2 + (0.15 if is_ti else 0.0)
    ph = np.exp(-1j * d[..., None] * k)                      # (A, V, F)
    gm = scale * np.einsum("v,avf->af", g, np.conj(ph))
    Q = np.einsum("af,avf->av", k[None, :] * meas, ph)
    w = np.imag(np.conj(g)[None, :] * Q)                     # (A, V)
    grad_d = -(e1 / n1[..., None] + e2 / n2[..., None])      # d(d)/dp
    ga = scale * np.einsum("av,avk->ak", w, grad_d)
    return gm, ga


@pytest.mark.parametrize("n_list", [1, 77, 129, None])
def test_backward_ragged_bistatic(n_list):
    """Voxel counts that are not multiples of the kernel's 128-voxel padding,
    a sample count (40) that leaves a short last block, two RX offsets and the
    TI range offset, against an FP64 NumPy backward."""
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
                                      is_ti=True)
    gm64, ga64 = _backward64(g, meas.astype(np.complex64).astype(np.complex128), ant,
                             rxo.astype(np.float32).astype(np.float64), v, sc.f0, sc.df,
                             scale, True)
    gm, ga = gm.cpu().numpy().astype(np.complex128), ga.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(gm - gm64) <= 1e-4 * np.linalg.norm(gm64)
    assert np.linalg.norm(ga - ga64) <= 1e-3 * np.linalg.norm(ga64)


@pytest.mark.parametrize("K", [16, 32])
@pytest.mark.parametrize("n_list", [77, None])
def test_backward_single_outputs(K, n_list):
    """The kernels for one gradient alone (the adjoint without the position
    term; the position gradient without the adjoint accumulators, at twice
    the waves) against the kernel for both and against the FP64 NumPy
    backward, on the ragged bistatic TI case."""
    sc = small_case(seed=6, n_ap=6, F=40, grid=(9, 7, 5))
    rng = np.random.default_rng(12)
    rxo = np.array([[0.0, 0.0, 0.0], [3e-3, 2e-3, 0.0]])
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
    args = (T(g, np.complex64), T(meas, np.complex64), T(ant, np.float32), xl, yl, zl,
            T(rxo, np.float32))
    kw = dict(f0=sc.f0, df=sc.df, scale=scale, vox_idx=idx, is_ti=True, K=K)
    gm, ga = ops.backproject_backward(*args, **kw)
    gm1, none_a = ops.backproject_backward(*args, need_ant=False, **kw)
    none_m, ga1 = ops.backproject_backward(*args, need_meas=False, **kw)
    assert none_a is None and none_m is None
    # the same sums per wave; only the chunking (occupancy) can differ
    assert (torch.linalg.norm(gm1 - gm) / torch.linalg.norm(gm)).item() <= 1e-6
    assert (torch.linalg.norm(ga1 - ga) / torch.linalg.norm(ga)).item() <= 1e-6
    gm64, ga64 = _backward64(g, meas.astype(np.complex64).astype(np.complex128), ant,
                             rxo.astype(np.float32).astype(np.float64), v, sc.f0, sc.df,
                             scale, True)
    gm1 = gm1.cpu().numpy().astype(np.complex128)
    ga1 = ga1.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(gm1 - gm64) <= 1e-4 * np.linalg.norm(gm64)
    assert np.linalg.norm(ga1 - ga64) <= 1e-3 * np.linalg.norm(ga64)


@pytest.mark.parametrize("entry", ["mmn_bp_backward", "mmn_bp_profile_backward"])
def test_backward_of_no_voxels_is_zero(entry):
    """An empty voxel list through the C ABI: both gradients are filled with
    zeros (by a fill kernel on the stream; the outputs start as ones)."""
    from mmnorm_amd import _lib
    sc = small_case(F=32, grid=(4, 4, 4))
    meas, a = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    A, F = meas.shape
    gm = torch.ones((A, F), dtype=torch.complex64, device=dev())
    ga = torch.ones((A, 3), dtype=torch.float32, device=dev())
    g1 = torch.zeros(1, dtype=torch.complex64, device=dev())      # not read
    idx1 = torch.zeros(1, dtype=torch.int32, device=dev())        # not read
    rxo = torch.zeros((1, 3), dtype=torch.float32, device=dev())
    common = (_lib.ptr(gm), _lib.ptr(ga), g1.data_ptr(), 1.0, xl.data_ptr(), yl.data_ptr(),
              zl.data_ptr(), idx1.data_ptr(), 0, a.data_ptr(), meas.data_ptr(),
              rxo.data_ptr(), float(sc.f0), float(sc.df), 4, 4, 4, A, 1, F, 0)
    if entry == "mmn_bp_backward":
        _lib.call(entry, *common, 32, None, 0, _lib.stream_handle(dev()))
    else:
        import ctypes
        bb = (ctypes.c_double * 6)(*ops.grid_bbox(xl, yl, zl))
        _lib.call(entry, *common, ctypes.addressof(bb), None, 0, _lib.stream_handle(dev()))
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(gm).item()) == 0 and int(torch.count_nonzero(ga).item()) == 0


@pytest.mark.parametrize("n_list", [77, None])
def test_position_gradient_with_64_sample_blocks(n_list):
    """The position-gradient kernel with 64-sample blocks (the default when
    only dL/dant is asked for; F = 150 leaves a short last block), and with 32
    and 16, against the FP64 NumPy backward on the bistatic TI case; K = 64
    with the adjoint is refused.  The blocks' real-coefficient phasor
    recurrence accumulates rounding over a block, so its error grows with K."""
    sc = small_case(seed=8, n_ap=6, F=150, grid=(9, 7, 5))
    rng = np.random.default_rng(13)
    rxo = np.array([[0.0, 0.0, 0.0], [3e-3, -2e-3, 0.0]])
    meas = np.repeat(sc.meas, 2, axis=0)[: 2 * len(sc.ant)]
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
    args = (T(g, np.complex64), T(meas, np.complex64), T(ant, np.float32), xl, yl, zl,
            T(rxo, np.float32))
    kw = dict(f0=sc.f0, df=sc.df, scale=scale, vox_idx=idx, is_ti=True)
    _, want = _backward64(g, meas.astype(np.complex64).astype(np.complex128), ant,
                          rxo.astype(np.float32).astype(np.float64), v, sc.f0, sc.df, scale,
                          True)
    err = {}
    for K in (16, 32, None):
        none_m, ga = ops.backproject_backward(*args, K=K, need_meas=False, **kw)
        assert none_m is None
        got = ga.cpu().numpy().astype(np.float64)
        err[K or 64] = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    print("position gradient rel. L2 error vs FP64 by block length:", err)
    assert all(e <= 1e-4 for e in err.values()), err
    with pytest.raises(ValueError, match="64"):
        ops.backproject_backward(*args, K=64, **kw)


def test_backward_workspace_smaller_than_the_plan():
    """The launch plans enough voxel chunks to keep each chunk's records in
    an XCD's L2 (here: 64 rows and 32,768 voxels, so many chunks of few
    voxels); given a quarter of the workspace the size query asks for, it
    uses fewer chunks and gives the same gradients (the chunk sums add in
    another grouping: to FP32 rounding); a workspace too small for one chunk
    is refused."""
    from mmnorm_amd import _lib
    sc = small_case(seed=3, F=64, grid=(32, 32, 32))
    meas, a = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    A, F = meas.shape
    V = xl.numel() * yl.numel() * zl.numel()
    rng = np.random.default_rng(4)
    g = T(rng.normal(size=V) + 1j * rng.normal(size=V), np.complex64)
    rxo = torch.zeros((1, 3), dtype=torch.float32, device=dev())
    full = int(_lib.load().mmn_bp_backward_workspace_size(A, F, V))
    outs = []
    for ws_bytes in (full, full // 4):
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev())
        gm = torch.empty((A, F), dtype=torch.complex64, device=dev())
        ga = torch.empty((A, 3), dtype=torch.float32, device=dev())
        _lib.call("mmn_bp_backward", gm.data_ptr(), ga.data_ptr(), g.data_ptr(), 1.0 / A,
                  xl.data_ptr(), yl.data_ptr(), zl.data_ptr(), None, 0, a.data_ptr(),
                  meas.data_ptr(), rxo.data_ptr(), float(sc.f0), float(sc.df), xl.numel(),
                  yl.numel(), zl.numel(), A, 1, F, 0, 32, ws.data_ptr(), ws_bytes,
                  _lib.stream_handle(dev()))
        outs.append((gm.cpu().numpy(), ga.cpu().numpy()))
    (m0, a0), (m1, a1) = outs
    assert np.linalg.norm(m1 - m0) <= 1e-6 * np.linalg.norm(m0)
    assert np.linalg.norm(a1 - a0) <= 1e-5 * np.linalg.norm(a0)
    ws = torch.empty(4096, dtype=torch.uint8, device=dev())
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("mmn_bp_backward", None, ga.data_ptr(), g.data_ptr(), 1.0, xl.data_ptr(),
                  yl.data_ptr(), zl.data_ptr(), None, 0, a.data_ptr(), meas.data_ptr(),
                  rxo.data_ptr(), float(sc.f0), float(sc.df), xl.numel(), yl.numel(),
                  zl.numel(), A, 1, F, 0, 32, ws.data_ptr(), 4096, _lib.stream_handle(dev()))


@pytest.mark.parametrize("variant", ["rec", "profile"])
def test_position_gradient_config4_geometry_vs_fd(oracle, variant):
    """Config 4 itself (40x40 aperture at 2 mm, 128 samples 77-81 GHz, 96^3
    voxels at 2 mm): dL/dp of L = sum_{v in S} w_v |I_v|^2 over a subset S of
    the 128 strongest and 128 random voxels, against FP64 central differences
    (h = 1e-6 m) on 64 random coordinates.  Tolerance from FP32: each
    coordinate's error at most 1e-5 of the sum of the magnitudes of its terms
    (tests/fd_subset.py), ten times the recurrence's measured relative error."""
    import fd_subset
    sc = synth.make_scene(4)
    ant = sc.ant.astype(np.float32).astype(np.float64)
    meas = T(sc.meas, np.complex64)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    a0 = T(ant, np.float32)
    img = ops.backproject(meas, a0, xl, yl, zl, f0=sc.f0, df=sc.df).reshape(-1)
    rng = np.random.default_rng(44)
    top = torch.topk(img.abs(), 128).indices.cpu().numpy()
    vidx = np.unique(np.concatenate([top, rng.choice(img.numel(), 128, replace=False)]))
    w = rng.uniform(size=len(vidx))
    scale = 1.0 / len(ant)
    a = a0.clone().requires_grad_(True)
    bb = ops.grid_bbox(xl, yl, zl)
    I = ops.backproject_ad(meas, a, xl, yl, zl, f0=sc.f0, df=sc.df, scale=scale,
                           vox_idx=T(vidx, np.int32), variant=variant, bbox=bb)
    (T(w, np.float32) * I.abs() ** 2).sum().backward()
    grad = a.grad.cpu().numpy().astype(np.float64)
    coords = [(int(rng.integers(len(ant))), int(rng.integers(3))) for _ in range(64)]
    fd, bound = fd_subset.fd_and_bound(oracle, sc, ant, vidx, w, scale, coords)
    got = np.array([grad[i, k] for (i, k) in coords])
    assert np.all(np.abs(got - fd) <= 1e-5 * bound), np.max(np.abs(got - fd) / bound)
    assert np.linalg.norm(got - fd) <= 1e-3 * np.linalg.norm(fd)
