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


def test_channel_synthesis_many_hits_and_none(oracle):
    """The synthesis over the dense hit list at its limits: more hits than its
    64 chunks of 512 (every chunk longer than the minimum) against the
    oracle, and a camera whose rays all miss: exactly zero.  The bound is the
    rays whose hit differs between the two marches (ray_matching on both
    sides, the same march at tau = 0): a ray that hits in one only moves a
    sample by at most 1, a hit moved by at most 1e-4 m by at most 0.7."""
    grid, bbox = sphere_scene()
    eye, rxd = [0.002, -0.001, 0.3], (0.0, 0.0, 1.0)
    W, H, F = 384, 384, 96
    f0, df = 77e9, 4e9 / (F - 1)
    gt = torch.from_numpy(grid).to(dev())
    got = ops.simulate_channel(gt, 0.0, W, H, 8.0, 8.0, bbox, eye, rxd, eye, VS, f0, df,
                               F).cpu().numpy()
    ref, dist = oracle.simulate_channel(grid, 0.0, W, H, 8.0, 8.0, bbox, eye, rxd, eye, VS, f0,
                                        df, F)
    nhit = np.count_nonzero(~np.isnan(dist))
    assert nhit > 64 * 512
    hit, vox, _ = ops.ray_matching(gt, W, H, 8.0, 8.0, bbox, eye, rxd)
    hit, vox = hit.reshape(-1, 3).cpu().numpy(), vox.reshape(-1, 3).cpu().numpy()
    ohit, ovox, _ = oracle.ray_matching(grid, W, H, 8.0, 8.0, bbox, eye, rxd)
    gm, om = vox[:, 0] != -1, ovox[:, 0] != -1
    both = gm & om
    moved = np.abs(hit[both] - ohit[both]).max(axis=1) if both.any() else np.zeros(0)
    assert moved.size == 0 or moved.max() <= 1e-4
    bound = 2.0 * np.count_nonzero(gm != om) + 0.7 * np.count_nonzero(moved > 0) + 1e-5 * nhit
    err = np.max(np.abs(got - ref))
    print(f"{nhit} hits, {np.count_nonzero(gm != om)} masks differ, "
          f"{np.count_nonzero(moved > 0)} hits moved: max error {err:.3g} (bound {bound:.3g}), "
          f"rel L2 {np.linalg.norm(got - ref) / np.linalg.norm(ref):.3g}")
    assert err <= bound
    # beside the box, looking away from it: no ray hits (a camera above the
    # box looking up is another matter -- the reference's first step takes
    # the box distance -1 plus 1e-5 as a hit at the eye, and so does the
    # oracle)
    eye2 = [0.3, 0.0, 0.04]
    _, dist2 = oracle.simulate_channel(grid, 0.0, W, H, 8.0, 8.0, bbox, eye2, (-1.0, 0.0, 0.0),
                                       eye2, VS, f0, df, F)
    assert np.all(np.isnan(dist2))
    away = ops.simulate_channel(gt, 0.0, W, H, 8.0, 8.0, bbox, eye2, (-1.0, 0.0, 0.0), eye2, VS,
                                f0, df, F)
    assert torch.count_nonzero(away) == 0


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


@pytest.mark.parametrize("F,N,lo,hi,hi_t", [(512, 512, 40, 90, 70),
                                             # twiddles without the LDS table, samples
                                             # read from global memory
                                             (1500, 4096, 300, 420, 400)])
def test_channel_loss_matches_numpy_fft(F, N, lo, hi, hi_t):
    rng = np.random.default_rng(3)
    B = 5
    sim = (rng.normal(size=(B, F)) + 1j * rng.normal(size=(B, F))).astype(np.complex64)
    tgt = (rng.normal(size=(B, F)) + 1j * rng.normal(size=(B, F))).astype(np.complex64)
    got = ops.channel_loss(torch.from_numpy(sim).to(dev()), torch.from_numpy(tgt).to(dev()),
                           N, lo, hi, hi_t).cpu().numpy()
    for b in range(B):
        a = np.fft.fft(sim[b].astype(np.complex128), n=N)
        t = np.fft.fft(tgt[b].astype(np.complex128), n=N)
        t[hi_t:] = 0
        a, t = a[lo:hi], t[lo:hi]
        a /= np.linalg.norm(a)
        t /= np.linalg.norm(t)
        want = np.sum(np.abs(np.abs(a) - np.abs(t)))
        assert abs(got[b] - want) <= 1e-5 * want
    # a channel against itself: exactly zero, as the reference's difference of
    # two normalised tensors
    same = ops.channel_loss(torch.from_numpy(sim).to(dev()), torch.from_numpy(sim).to(dev()),
                            N, lo, hi, hi).cpu().numpy()
    assert np.all(same == 0.0)


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


def two_component_grid():
    """Two RSDF-like components of different radii, component-major."""
    g1, bbox = sphere_scene()
    g2, _ = sphere_scene(radius=0.012)
    return np.stack([g1, g2]), bbox


def test_simulate_channels_batch_equals_single_calls(monkeypatch):
    """Each simulation of a batch -- its own antenna, direction, offset,
    component, TI flag shared -- equals mmn_simulate_channel bit for bit, also
    when the wrapper splits the batch over several launches."""
    gcm, bbox = two_component_grid()
    gt = torch.from_numpy(gcm).to(dev())
    rng = np.random.default_rng(5)
    n = 7
    comp = rng.integers(0, 2, n)
    taus = rng.uniform(-0.003, 0.003, n).astype(np.float32)
    eyes = np.stack([rng.uniform(-0.01, 0.01, n), rng.uniform(-0.01, 0.01, n),
                     rng.uniform(0.25, 0.3, n)], 1)
    dirs = np.stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), np.ones(n)], 1)
    W, H, F = 40, 36, 96
    f0, df = 77e9, 4e9 / (F - 1)
    for is_ti in (False, True):
        one = torch.stack([ops.simulate_channel(
            gt[int(comp[s])].contiguous(), float(taus[s]), W, H, 8.0, 7.0, bbox,
            np.float32(eyes[s]).tolist(), np.float32(dirs[s]).tolist(), eyes[s].tolist(), VS,
            f0, df, F, is_ti=is_ti) for s in range(n)])
        assert torch.count_nonzero(one) > 0
        got = ops.simulate_channels(gt, comp, taus, eyes, dirs, eyes, W, H, 8.0, 7.0, bbox, VS,
                                    f0, df, F, is_ti=is_ti)
        assert torch.equal(got, one)
        monkeypatch.setattr(ops, "MAX_BATCH_RAYS", 3 * W * H)   # three launches of 3, 3, 1
        split = ops.simulate_channels(gt, comp, taus, eyes, dirs, eyes, W, H, 8.0, 7.0, bbox,
                                      VS, f0, df, F, is_ti=is_ti)
        monkeypatch.undo()
        assert torch.equal(split, one)


def test_simulate_channels_rejects_bad_components():
    gcm, bbox = two_component_grid()
    gt = torch.from_numpy(gcm).to(dev())
    args = ([0.0], [[0, 0, 0.3]], [[0, 0, 1]], [[0, 0, 0.3]], 8, 8, 8.0, 8.0, bbox, VS, 77e9,
            1e8, 16)
    with pytest.raises(ValueError):
        ops.simulate_channels(gt, [2], *args)
    with pytest.raises(ValueError):
        ops.simulate_channels(gt, [-1], *args)


def test_channel_batch_equals_per_antenna_dropin():
    """isosurface's batch of one offset over several antennas and both
    components (one offset clipped) equals _simulate_channels per antenna."""
    gcm, bbox = two_component_grid()
    g_last = torch.from_numpy(np.ascontiguousarray(np.moveaxis(gcm, 0, -1))).to(dev())
    F = 64
    wl = torch.from_numpy(synth.C0 / (77e9 + 4e9 / (F - 1) * np.arange(F)))
    locs = torch.tensor([[0.001, 0.0, 0.3], [-0.004, 0.003, 0.28], [0.006, -0.002, 0.31]],
                        dtype=torch.float64)
    dirs = torch.tensor([[0.0, 0.0, 1.0], [0.05, 0.0, 1.0], [0.0, -0.04, 1.0]])
    bounds = np.array([[-0.004, 0.004], [-0.004, 0.004]])
    taus = torch.tensor([0.0013, 0.5])
    want = torch.stack([isosurface._simulate_channels(
        bbox, None, VS, *gcm.shape[1:], 24, 20, 8.0, 8.0, g_last, locs[r], dirs[r], wl, taus,
        bounds, True) for r in range(3)])
    got = isosurface._simulate_channel_batch(bbox, VS, 24, 20, 8.0, 8.0,
                                             torch.from_numpy(gcm).to(dev()), locs, dirs, wl,
                                             taus, bounds, True)
    assert torch.count_nonzero(want) > 0
    assert torch.equal(got, want)


def test_search_taus_matches_per_call_loop_and_finds_generating_offset():
    """search_taus (isosurface_optimization.py:454-521) against the loop it
    replaces -- _simulate_channels per (offset, antenna) and batch_loss --
    bit for bit, on targets simulated at one offset of the grid: the loss is
    zero there and larger everywhere else."""
    grid, bbox = sphere_scene()
    sdf = torch.from_numpy(grid[..., None]).to(dev())
    F = 64
    wl = torch.from_numpy(synth.C0 / (77e9 + 4e9 / (F - 1) * np.arange(F)))
    rng = np.random.default_rng(11)
    n_ant = 24
    locs = torch.from_numpy(np.stack([rng.uniform(-0.012, 0.012, n_ant),
                                      rng.uniform(-0.012, 0.012, n_ant),
                                      np.full(n_ant, 0.3)], 1))
    dirs = torch.zeros(n_ant, 3)
    dirs[:, 2] = 1.0
    bounds = np.array([[-0.006, 0.006]])
    step = float(wl[0]) / 2
    tau_axis = torch.arange(-0.006, 0.006, step, dtype=torch.float32)
    t_star = tau_axis[2:3]
    chans = torch.stack([isosurface._simulate_channels(
        bbox, None, VS, *grid.shape, 32, 32, 10.0, 10.0, sdf, locs[r], dirs[r], wl, t_star,
        bounds, False) for r in range(n_ant)])
    kw = dict(is_real_data=False, width=32, height=32, width_fov=10.0, height_fov=10.0,
              num_rx=10, skip=0, min_dist=0.2, max_dist=0.4)
    taus, tau_grid, losses = isosurface.search_taus(
        sdf, locs, dirs, chans, bbox, wl, VS, bounds, F,
        generator=torch.Generator().manual_seed(3), **kw)
    # the loop of the reference, on the same antennas
    ch_f, loc_f, dir_f = isosurface._filter_channels(chans, locs, dirs, False)
    assert len(ch_f) >= 10
    order = torch.randperm(len(loc_f), generator=torch.Generator().manual_seed(3))[:10]
    want = torch.zeros(len(tau_grid), dtype=torch.float32)
    for i in range(len(tau_grid)):
        sims = torch.stack([isosurface._simulate_channels(
            bbox, None, VS, *grid.shape, 32, 32, 10.0, 10.0, sdf, loc_f[r], dir_f[r], wl,
            tau_grid[i], bounds, False) for r in order.tolist()])
        want[i] = isosurface.batch_loss(sims, ch_f[order].to(dev()), False).sum().cpu()
    assert torch.equal(losses, want)
    assert torch.equal(tau_grid[:, 0], tau_axis)
    assert float(losses[2]) == 0.0
    assert float(torch.min(torch.cat([losses[:2], losses[3:]]))) > 0.0


@pytest.mark.parametrize("res", [(40, 40, 40), (30, 50, 20)])
def test_integer_corner_indices_equal_fp32_indices(monkeypatch, res):
    """The sphere trace's corner indices in integers (grids up to 2^24
    voxels) against the reference's FP32 index arithmetic
    (MMN_RAY_FP32_INDEX=1): the same hits, voxels and channels, bit for bit,
    also on a grid whose x extent exceeds what the y-derived voxel size
    spans (indices past the axis wrap or clamp the same)."""
    grid, bbox = sphere_scene(res=res)
    if res[0] != res[1]:
        bbox = list(bbox)
        bbox[3] += 0.01          # x extent beyond (res_x - 1) voxels of the y size
    gt = torch.from_numpy(grid).to(dev())
    eye, rxd = [0.003, -0.002, 0.3], (0.02, -0.01, 1.0)
    W, H, F = 96, 80, 64

    def run():
        rm = ops.ray_matching(gt, W, H, 12.0, 10.0, bbox, eye, rxd)
        ch = ops.simulate_channel(gt, 0.0013, W, H, 12.0, 10.0, bbox, eye, rxd, eye, VS, 77e9,
                                  4e9 / (F - 1), F, is_ti=True)
        return [t.cpu() for t in rm] + [ch.cpu()]

    fast = run()
    monkeypatch.setenv("MMN_RAY_FP32_INDEX", "1")
    ref = run()
    assert (fast[1][..., 0] != -1).sum() > 300        # rays hit the sphere
    for a, b in zip(fast, ref):
        assert torch.equal(a, b)


def test_prefixed_channels_equal_full_march(monkeypatch):
    """The offset-independent prefix (mmn_ray_prefix) and the resumed march of
    the live rays (mmn_simulate_channels_prefixed) give mmn_simulate_channels'
    channels bit for bit for offsets >= tau_lo: antennas above, inside and
    within a voxel of the box (the first step's box-distance hit), a band of
    corners just below the undefined threshold (undefined for some offsets of
    the range and not for others), several launch groups."""
    gcm, bbox = two_component_grid()
    gcm = gcm.copy()
    band = (gcm[0] > 0.004) & (gcm[0] < 0.0055)
    gcm[0][band] = -40.0 + 0.0015          # undefined iff tau > 0.0015
    gt = torch.from_numpy(gcm).to(dev())
    eyes = np.array([[0.002, -0.001, 0.3], [0.01, 0.004, 0.25], [-0.006, 0.003, 0.07],
                     [0.0, 0.002, bbox[5] + 0.0005], [0.004, -0.008, 0.28]])
    dirs = np.array([[0.0, 0.0, 1.0], [0.05, -0.02, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0],
                     [-0.03, 0.04, 1.0]])
    n = len(eyes)
    comp = np.array([0, 1, 0, 1, 0])
    W, H, F = 48, 40, 96
    f0, df = 77e9, 4e9 / (F - 1)
    monkeypatch.setattr(ops, "MAX_BATCH_RAYS", 2 * W * H)     # groups of 2, 2, 1
    tau_lo = -0.003
    pre = ops.ray_prefix(gt, comp, eyes, dirs, W, H, 12.0, 10.0, bbox, tau_lo)
    assert len(pre.groups) == 3
    assert 0 < pre.live_rays < n * W * H
    for taus in (np.full(n, tau_lo), np.array([0.001, 0.0, 0.002, -0.001, 0.0016]),
                 np.full(n, 0.003)):
        taus = taus.astype(np.float32)
        want = ops.simulate_channels(gt, comp, taus, eyes, dirs, eyes, W, H, 12.0, 10.0, bbox,
                                     VS, f0, df, F, is_ti=True)
        got = ops.simulate_channels_prefixed(pre, taus, eyes, VS, f0, df, F, is_ti=True)
        assert torch.count_nonzero(want) > 0
        assert torch.equal(got, want)
    with pytest.raises(ValueError):
        ops.simulate_channels_prefixed(pre, np.full(n, -0.004, np.float32), eyes, VS, f0, df,
                                       F)
