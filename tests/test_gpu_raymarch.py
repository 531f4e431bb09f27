"""Ray marching and channel simulation on the GPU against the oracle's
restatement of ray_tracing_kernel.cu / isosurface_optimization.py, closed
forms, and the drop-ins of mmnorm_amd.isosurface."""
import numpy as np
import pytest
import torch

from mmnorm_amd import isosurface, ops, synth

pytestmark = pytest.mark.gpu

VS = 2e-3


def dev():
    return torch.device("cuda", 0)


def sphere_scene(res=(40, 40, 40), radius=0.018, undef_outside=0.006):
    """RSDF-like grid: signed distance to a sphere, undefined (-50) farther
    than ``undef_outside`` from it; the grid box and its corner coordinates."""
    bmin = np.array([-0.039, -0.039, 0.0])
    axes = [bmin[a] + VS * np.arange(res[a]) for a in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    c = np.array([0.0, 0.0, 0.04])
    sd = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - radius
    grid = sd.astype(np.float32)
    grid[np.abs(sd) > undef_outside] = -50.0
    bbox = list(bmin) + list(bmin + VS * (np.array(res) - 1))
    return grid, bbox


def test_ray_matching_matches_oracle(oracle):
    grid, bbox = sphere_scene()
    eye, rxd = [0.002, -0.001, 0.3], (0.0, 0.0, 1.0)
    W, H = 64, 48
    hit, vox, d = ops.ray_matching(torch.from_numpy(grid).to(dev()), W, H, 8.0, 6.0, bbox, eye,
                                   rxd)
    hit, vox, d = (t.reshape(-1, 3).cpu().numpy() for t in (hit, vox, d))
    ohit, ovox, od = oracle.ray_matching(grid, W, H, 8.0, 6.0, bbox, eye, rxd)
    assert np.max(np.abs(d - od)) <= 2e-7
    gm, om = vox[:, 0] != -1, ovox[:, 0] != -1
    assert gm.sum() > 500                       # the sphere is in view
    assert np.mean(gm == om) >= 0.995
    both = gm & om
    assert np.median(np.abs(hit[both] - ohit[both])) == 0.0
    assert np.max(np.abs(hit[both] - ohit[both])) <= 1e-4


def test_simulate_channel_matches_oracle(oracle):
    grid, bbox = sphere_scene()
    eye, rxd = [0.002, -0.001, 0.3], (0.0, 0.0, 1.0)
    W, H, F = 64, 48, 128
    f0, df = 77e9, 4e9 / (F - 1)
    for tau, is_ti in ((0.0, False), (-0.0021, True)):
        got = ops.simulate_channel(torch.from_numpy(grid).to(dev()), tau, W, H, 8.0, 6.0, bbox,
                                   eye, rxd, eye, VS, f0, df, F, is_ti=is_ti).cpu().numpy()
        ref, dist = oracle.simulate_channel(grid, tau, W, H, 8.0, 6.0, bbox, eye, rxd, eye, VS,
                                            f0, df, F, is_ti=is_ti)
        nhit = np.count_nonzero(~np.isnan(dist))
        assert nhit > 500
        # a ray that hits in one and not the other moves the sum by at most 1
        # per sample; the recurrence itself is good to ~1e-6 per term
        err = np.max(np.abs(got - ref))
        assert err <= 2.0 + 1e-5 * nhit, (err, nhit)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-2


def test_single_vertical_ray_channel_closed_form():
    """One ray straight down onto the plane z = 0.0311: d = 2 (eye_z - z0)."""
    res = (16, 16, 32)
    bmin = np.array([-0.015, -0.015, 0.0])
    bbox = list(bmin) + list(bmin + VS * (np.array(res) - 1))
    z = bmin[2] + VS * np.arange(res[2])
    z0 = 0.0311
    grid = np.broadcast_to((z - z0).astype(np.float32), res).copy()
    eye = [0.0013, -0.0021, 0.25]
    F, f0, df = 64, 77e9, 4e9 / 63
    ch = ops.simulate_channel(torch.from_numpy(grid).to(dev()), 0.0, 1, 1, 40.0, 40.0, bbox,
                              eye, (0, 0, 1), eye, VS, f0, df, F).cpu().numpy()
    d = 2 * (eye[2] - z0)
    want = np.exp(2j * np.pi * d * (f0 + df * np.arange(F)) / synth.C0)
    assert np.max(np.abs(ch - want)) <= 2e-3   # FP32 hit position (~1e-7 m) and recurrence


def test_channel_loss_matches_numpy_fft():
    rng = np.random.default_rng(3)
    B, F, N, lo, hi = 5, 512, 512, 40, 90
    sim = (rng.normal(size=(B, F)) + 1j * rng.normal(size=(B, F))).astype(np.complex64)
    tgt = (rng.normal(size=(B, F)) + 1j * rng.normal(size=(B, F))).astype(np.complex64)
    got = ops.channel_loss(torch.from_numpy(sim).to(dev()), torch.from_numpy(tgt).to(dev()),
                           N, lo, hi, 70).cpu().numpy()
    for b in range(B):
        a = np.fft.fft(sim[b].astype(np.complex128), n=N)
        t = np.fft.fft(tgt[b].astype(np.complex128), n=N)
        t[70:] = 0
        a, t = a[lo:hi], t[lo:hi]
        a /= np.linalg.norm(a)
        t /= np.linalg.norm(t)
        want = np.sum(np.abs(np.abs(a) - np.abs(t)))
        assert abs(got[b] - want) <= 1e-5 * want


def test_ray_matching_dropin_shapes_and_values():
    grid, bbox = sphere_scene()
    g = torch.from_numpy(grid).to(dev())
    W, H = 32, 32
    out = isosurface.ray_matching(torch.zeros(W, H, 3, device=dev()), torch.zeros(W, H, device=dev()),
                                  g, W, H, 8.0, 8.0, *bbox, *grid.shape, 0.0, 0.0, 0.3, 0.0, 0.0,
                                  1.0)
    hit, vox, d = ops.ray_matching(g, W, H, 8.0, 8.0, bbox, [0.0, 0.0, 0.3], [0.0, 0.0, 1.0])
    assert [tuple(t.shape) for t in out] == [(W, H, 3)] * 3
    assert torch.equal(out[0], hit) and torch.equal(out[1], vox) and torch.equal(out[2], d)


def test_simulate_channels_dropin_sums_components_and_clips_tau():
    grid, bbox = sphere_scene()
    g2 = torch.from_numpy(np.stack([grid, grid], -1)).to(dev())
    F = 64
    f = 77e9 + 4e9 / (F - 1) * np.arange(F)
    wl = torch.from_numpy(synth.C0 / f)
    rx = torch.tensor([0.001, 0.0, 0.3], dtype=torch.float64)
    rxd = torch.tensor([0.0, 0.0, 1.0])
    bounds = torch.tensor([[-0.004, 0.004], [-0.004, 0.004]])
    taus = torch.tensor([0.001, 0.5])          # the second is clipped to 0.004
    both = isosurface._simulate_channels(bbox, None, VS, *grid.shape, 24, 24, 8.0, 8.0, g2, rx,
                                         rxd, wl, taus, bounds, False)
    one = [ops.simulate_channel(torch.from_numpy(grid).to(dev()), t, 24, 24, 8.0, 8.0, bbox,
                                [0.001, 0.0, 0.3], [0, 0, 1], rx.tolist(), VS, f[0], f[1] - f[0],
                                F) for t in (0.001, 0.004)]
    assert torch.allclose(both, one[0] + one[1], atol=1e-3)
