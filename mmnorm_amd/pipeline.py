"""End-to-end synthetic pipeline: forward with the fused vote, RSDF BFS, then
isosurface optimisation of the antenna positions through the backward kernel.

    python -m mmnorm_amd.pipeline --config 5 --steps 50 [--variant profile] [--out DIR]

1. Scene (``synth``), true antenna positions p*, start positions
   p0 = p* + N(0, sigma) (sigma = 0.3 mm by default).
2. Full-grid forward at p0 (recurrence kernel, vote moment fused).
3. Normals on voxels above 5 % of the image maximum (py_sn_est_gpu.py:326),
   unit-normalised (image_processor.py:396); the Li binary image of |I|^2
   removes the rest (sdf.py:276-279); 6-connected components on the
   GPU and the reference's size rule (sdf.py:88-119); per-component seeds
   (sdf.py:182-191); GPU RSDF BFS.
4. Isosurface band B = {reached voxels with |rsdf| <= tau}.  The loss is the
   negative image energy on the band, L(p) = -sum_{v in B} |I_v(p)|^2 / E0,
   with I = backprojection / A over the band voxels only, minimised with Adam
   over the range (z) coordinate of every antenna position: for a planar
   aperture the in-plane coordinates move the round-trip distance by only
   the sine of the look angle, so an energy objective barely constrains them
   and Adam's normalised steps would random-walk them.  Every step is one
   forward and one backward kernel launch over the band (phases recomputed,
   nothing stored per voxel x row x sample).
   ``objective="sharpness"`` maximises sum |I|^4 / (sum |I|^2)^2 instead.
   With one free range per antenna both objectives can be pushed past their
   value at the true positions (noise is focused too), so the step size is
   kept small: at config 5 the energy objective with lr = 2e-6 lowers the
   range error steadily (0.301 -> 0.254 mm in 50 steps) while larger steps
   or the sharpness objective overshoot.  ``variant="profile"`` runs the
   range-profile forward and the grouped-series backward (the same
   trajectory, about 5x faster steps).
The reference has no gradient loop (its isosurface stage is a grid search,
isosurface_optimization.py:462-521); this loss and optimiser are defined here.
Progress is printed per step; ``--out`` also writes a JSON record and an
optimiser checkpoint every ``--ckpt-every`` steps (resumed with ``--resume``).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from . import ops, sdf, synth


C0_M = 2.99792458e8


def log(msg):
    print(f"[pipeline] {msg}", flush=True)


def _objective(I, kind):
    p = I.abs() ** 2
    if kind == "energy":
        return p.sum()
    if kind == "sharpness":  # sum |I|^4 / (sum |I|^2)^2: scale-free image sharpness
        return (p * p).sum() / p.sum() ** 2
    raise ValueError(f"unknown objective {kind!r}")


def tau_search_stage(sc, p0, rsdf, labeled, valid, coords, rays, fov, num_rx, seed, rec):
    """The RSDF consumer: the isosurface offset search of
    isosurface_optimization.py:454-521 on the synthetic scene, with every
    antenna's measured row as its channel, then the distance from the chosen
    isosurface's points to the true surface (the scatterers)."""
    from scipy.spatial import cKDTree
    from . import isosurface
    dev = rsdf.device
    comps, bounds = isosurface.prepare_components(rsdf.cpu().numpy(), None, labeled, valid)
    F = sc.cfg.n_freq
    wl = C0_M / (sc.f0 + sc.df * np.arange(F))
    bbox = [sc.x_locs[0], sc.y_locs[0], sc.z_locs[0], sc.x_locs[-1], sc.y_locs[-1],
            sc.z_locs[-1]]
    rx_locs = torch.from_numpy(np.asarray(p0, np.float64))
    rx_dirs = torch.zeros(len(p0), 3)
    rx_dirs[:, 2] = 1.0
    chans = torch.from_numpy(np.ascontiguousarray(sc.meas)).to(dev)
    torch.cuda.synchronize()
    t = time.perf_counter()
    taus, grid, losses = isosurface.search_taus(
        comps, rx_locs, rx_dirs, chans, bbox, wl, float(sc.z_locs[1] - sc.z_locs[0]), bounds, F,
        is_real_data=False, width=rays, height=rays, width_fov=fov, height_fov=fov,
        num_rx=num_rx, generator=torch.Generator().manual_seed(seed), skip=0, min_dist=0.2,
        max_dist=0.4)
    torch.cuda.synchronize()
    rec["tau_search_s"] = time.perf_counter() - t
    n_sim = int(np.prod(losses.shape)) * num_rx * comps.shape[-1]
    rec["tau_search_sims"] = n_sim
    tree = cKDTree(sc.pos)
    sdf_np = comps.cpu().numpy()

    def surf_err(tv):
        pts = isosurface._get_final_point_cloud(sdf_np, np.atleast_1d(tv), coords)
        return (float(np.median(tree.query(pts)[0])) if len(pts) else float("nan")), len(pts)

    flat = grid.reshape(-1, grid.shape[-1]).numpy()
    errs = [surf_err(tv)[0] for tv in flat]
    chosen_err, npts = surf_err(taus)
    rec.update({"taus": np.atleast_1d(taus).tolist(), "tau_grid_size": len(flat),
                "chosen_surface_err_mm": chosen_err * 1e3, "chosen_points": npts,
                "best_grid_surface_err_mm": float(np.nanmin(errs)) * 1e3,
                "tau0_surface_err_mm": surf_err(np.zeros(grid.shape[-1]))[0] * 1e3})
    log(f"tau search: {len(flat)} offsets x {num_rx} antennas x {rays}^2 rays in "
        f"{rec['tau_search_s']:.2f} s ({rec['tau_search_s'] / n_sim * 1e3:.2f} ms per "
        f"simulated channel); tau = {rec['taus']}, median distance of its {npts} points to the "
        f"true surface {rec['chosen_surface_err_mm']:.2f} mm (tau = 0: "
        f"{rec['tau0_surface_err_mm']:.2f} mm, best offset on the grid: "
        f"{rec['best_grid_surface_err_mm']:.2f} mm)")


def prepare(cid=5, sigma=3e-4, K=64, seed=7, variant="rec", repeat_prep=0, rec=None):
    """Steps 1-3 for configuration `cid`: the scene, the perturbed start
    positions p0, the full-grid forward at p0 with the vote moment, the unit
    normals on the 5 % filter without the voxels below the Li threshold of
    |I|^2, the 6-connected labels and the size rule, the start voxels and
    the RSDF BFS.  Returns a dict of the scene, the device fields and the
    host seeds; `rec` (a dict) receives the timings."""
    rec = {} if rec is None else rec
    dev = torch.device("cuda", torch.cuda.current_device())
    t0 = time.perf_counter()
    sc = synth.make_scene(cid, device=dev)
    cfg = sc.cfg
    log(f"config {cid}: scene ready in {time.perf_counter() - t0:.1f} s "
        f"(A={cfg.rows}, F={cfg.n_freq}, grid={cfg.grid})")
    T = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(dev)
    meas = T(sc.meas, np.complex64)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    rng = np.random.default_rng(seed)
    p_true = sc.ant.astype(np.float32)
    p0 = (sc.ant + rng.normal(scale=sigma, size=sc.ant.shape)).astype(np.float32)
    bb = ops.grid_bbox(xl, yl, zl)

    # 2. full-grid forward with the fused vote moment
    torch.cuda.synchronize()
    t = time.perf_counter()
    img, mom = ops.backproject(meas, T(p0, np.float32), xl, yl, zl, f0=sc.f0, df=sc.df,
                               K=K, moment=True, variant=variant, bbox=bb)
    torch.cuda.synchronize()
    rec["forward_full_s"] = time.perf_counter() - t
    log(f"full-grid forward: {rec['forward_full_s']:.2f} s "
        f"({cfg.work / rec['forward_full_s'] / 1e9:.0f} G voxel*row*sample/s)")

    # 3. normals, components, seeds, BFS
    t = time.perf_counter()
    mag = img.abs().reshape(-1)
    vidx = torch.nonzero(mag > 0.05 * mag.max()).reshape(-1).to(torch.int32)
    n = ops.sn_from_moment(img.reshape(-1), mom.reshape(-1, 3)[vidx.long()].contiguous(),
                           vidx).double()
    n = n / torch.linalg.norm(n, dim=1, keepdim=True)
    normals = torch.full((img.numel(), 3), float("nan"), dtype=torch.float64, device=dev)
    normals[vidx.long()] = n
    normals = normals.reshape(*img.shape, 3)
    # binary image (sdf.py:276-279): voxels below the Li threshold of |I|^2
    # lose their normals (the first exponent passes the volume test here:
    # the whole grid spans less than 0.03 m^3)
    thr, li_iters, _, below = ops.li_threshold(img.contiguous(), 2, below=True)
    normals[below] = float("nan")
    rec["li_threshold_iters"] = li_iters
    dx = float(sc.x_locs[1] - sc.x_locs[0])
    torch.cuda.synchronize()
    tc = time.perf_counter()
    lab_t, areas = sdf.find_components_gpu(normals)   # GPU 6-connected labels + 1
    valid = sdf.size_rule(areas, dx)
    torch.cuda.synchronize()
    rec["ccl_s"] = time.perf_counter() - tc
    # start voxels: the whole-grid masks on the GPU, each component's mean
    # and nearest voxel on the host (NumPy's sequential sum)
    seeds = sdf.component_seeds_gpu(normals, lab_t, valid, sc.x_locs, sc.y_locs, sc.z_locs)
    torch.cuda.synchronize()
    rec["seeds_s"] = time.perf_counter() - tc - rec["ccl_s"]
    tb = time.perf_counter()
    rsdf, hops = ops.rsdf_bfs(normals, lab_t, T(seeds, np.int32), dx)
    torch.cuda.synchronize()
    rec["bfs_s"] = time.perf_counter() - tb
    rec["prep_s"] = time.perf_counter() - t
    if repeat_prep > 0:
        # the labelling, the start voxels and the BFS again on the same fields:
        # the times above include each stage's first use in the process
        warm = {"ccl": [], "seeds": [], "bfs": []}
        seeds_t = T(seeds, np.int32)
        for _ in range(repeat_prep):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            sdf.find_components_gpu(normals)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            sdf.component_seeds_gpu(normals, lab_t, valid, sc.x_locs, sc.y_locs, sc.z_locs)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            ops.rsdf_bfs(normals, lab_t, seeds_t, dx)
            torch.cuda.synchronize()
            t4 = time.perf_counter()
            warm["ccl"].append(t2 - t1)
            warm["seeds"].append(t3 - t2)
            warm["bfs"].append(t4 - t3)
        for k, v in warm.items():
            rec[f"{k}_warm_ms"] = float(np.median(v)) * 1e3
        log(f"again ({repeat_prep} times, median): CCL {rec['ccl_warm_ms']:.2f} ms, "
            f"seeds {rec['seeds_warm_ms']:.2f} ms, BFS {rec['bfs_warm_ms']:.2f} ms")
    return {"sc": sc, "meas": meas, "axes": (xl, yl, zl), "bbox": bb, "p_true": p_true,
            "p0": p0, "img": img, "mom": mom, "vidx": vidx, "normals": normals,
            "labels": lab_t, "valid": valid, "seeds": seeds, "rsdf": rsdf, "hops": hops,
            "dx": dx, "t0": t0}


def run(cid=5, steps=50, sigma=3e-4, tau_vox=1.5, lr=2e-6, K=64, out=None,
        ckpt_every=10, resume=False, seed=7, objective="energy", tau_search=False, rays=512,
        fov=20.0, num_rx=50, variant="rec", repeat_prep=0):
    rec = {"config": cid, "sigma_m": sigma, "steps": steps, "K": K, "variant": variant}
    st = prepare(cid, sigma, K, seed, variant, repeat_prep, rec)
    sc, meas, (xl, yl, zl), bb = st["sc"], st["meas"], st["axes"], st["bbox"]
    p_true, p0, img, vidx, normals = st["p_true"], st["p0"], st["img"], st["vidx"], st["normals"]
    lab_t, valid, seeds, rsdf, hops, dx = (st["labels"], st["valid"], st["seeds"], st["rsdf"],
                                           st["hops"], st["dx"])
    t0 = st["t0"]
    cfg = sc.cfg
    A = cfg.rows
    dev = meas.device
    T = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(dev)
    reached = hops >= 0
    band = reached & (rsdf.abs() <= tau_vox * dx)
    bidx = torch.nonzero(band.reshape(-1)).reshape(-1).to(torch.int32)
    rec.update({"valid_voxels": int(vidx.numel()), "components": int(len(seeds)),
                "reached": int(reached.sum().item()), "levels": int(hops.max().item()) + 1,
                "band_voxels": int(bidx.numel())})
    log(f"vote + CCL + BFS: {rec['prep_s']:.2f} s (CCL {rec['ccl_s'] * 1e3:.1f} ms, "
        f"seeds {rec['seeds_s'] * 1e3:.1f} ms, BFS {rec['bfs_s'] * 1e3:.1f} ms, "
        f"{rec['components']} components, {rec['reached']} voxels reached, "
        f"{rec['levels']} levels); band {rec['band_voxels']} voxels")
    if bidx.numel() == 0:
        raise RuntimeError("empty isosurface band")
    if out or tau_search:
        # host copies for the stage file and the tau search (sdf.py:278-279)
        nrm_np = normals.cpu().numpy()
        labeled = lab_t.cpu().numpy()
        m1, m2, m3 = np.meshgrid(sc.y_locs, sc.x_locs, sc.z_locs)
        coords = np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1)
    if out:
        # the reference's stage file after the RSDF (sdf.py:226-253)
        from . import formats
        formats.save_sdf_file(os.path.join(out, f"sdf_config{cid}.npz"), formats.sdf_record(
            (img / A).cpu().numpy(), sc.x_locs, sc.y_locs, sc.z_locs, p0, sc.meas[:, :, None],
            nrm_np, rsdf.cpu().numpy(), labeled, valid, coords, sc.f0,
            sc.f0 + sc.df * cfg.n_freq, cfg.n_freq))
    if tau_search:
        tau_search_stage(sc, p0, rsdf, labeled, valid, coords, rays, fov, num_rx, seed, rec)

    # 4. optimisation over the band
    pos = T(p0, np.float32).requires_grad_(True)
    opt = torch.optim.Adam([pos], lr=lr)
    start = 0
    ck = os.path.join(out, f"pipeline_config{cid}.pt") if out else None
    if resume and ck and os.path.exists(ck):
        state = torch.load(ck, weights_only=True)
        with torch.no_grad():
            pos.copy_(state["pos"])
        opt.load_state_dict(state["opt"])
        start = int(state["step"])
        log(f"resumed at step {start}")
    def objective_at(p):
        with torch.no_grad():
            return _objective(ops.backproject(meas, p, xl, yl, zl, f0=sc.f0, df=sc.df, K=K,
                                              vox_idx=bidx, variant=variant, bbox=bb) / A,
                              objective).item()

    e0 = objective_at(pos.detach())
    hist = []
    pt = torch.from_numpy(p_true).to(dev)

    def pos_err(axis=None):
        d = (pos.detach() - pt)
        d = d - d.mean(0, keepdim=True)  # a common shift is not observable
        if axis is not None:
            return float(torch.sqrt((d[:, axis] ** 2).mean()).item())
        return float(torch.sqrt((d ** 2).sum(1).mean()).item())

    # a range error that is affine over the aperture plane only shifts the
    # image, which an energy objective cannot see: report the residual too
    Xa = torch.stack([pt[:, 0], pt[:, 1], torch.ones_like(pt[:, 0])], 1).double()

    def range_err_affine_removed():
        d = (pos.detach() - pt)[:, 2].double()
        fit = torch.linalg.lstsq(Xa, d[:, None]).solution
        return float(torch.sqrt(((d - (Xa @ fit)[:, 0]) ** 2).mean()).item())

    e_true = objective_at(pt)
    rec["objective"] = objective
    rec["objective_true_over_start"] = e_true / e0
    rec["start_rms_z_err_mm"] = pos_err(2) * 1e3
    rec["start_rms_z_err_affine_removed_mm"] = range_err_affine_removed() * 1e3
    log(f"start: band {objective} {e0:.4e} (true positions: {e_true / e0:.4f} x), rms position "
        f"error {pos_err() * 1e3:.3f} mm (range axis {pos_err(2) * 1e3:.3f} mm, "
        f"{rec['start_rms_z_err_affine_removed_mm']:.3f} mm without its affine part)")
    step_times = []
    for it in range(start, steps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        opt.zero_grad(set_to_none=True)
        I = ops.backproject_ad(meas, pos, xl, yl, zl, f0=sc.f0, df=sc.df, scale=1.0 / A,
                               K=K, vox_idx=bidx, variant=variant, bbox=bb)
        loss = -_objective(I, objective) / e0
        loss.backward()
        with torch.no_grad():
            pos.grad[:, :2] = 0  # range coordinate only (see module docstring)
        opt.step()
        torch.cuda.synchronize()
        step_times.append(time.perf_counter() - ts)
        hist.append({"step": it + 1, "loss": float(loss.item()), "rms_err_mm": pos_err() * 1e3,
                     "rms_z_err_mm": pos_err(2) * 1e3,
                     "rms_z_err_affine_removed_mm": range_err_affine_removed() * 1e3,
                     "s": step_times[-1]})
        log(f"step {it + 1}/{steps}: loss {hist[-1]['loss']:.6f}, rms position error "
            f"{hist[-1]['rms_err_mm']:.4f} mm (range axis {hist[-1]['rms_z_err_mm']:.4f} mm, "
            f"{hist[-1]['rms_z_err_affine_removed_mm']:.4f} mm non-affine), "
            f"{step_times[-1]:.2f} s")
        if ck and ((it + 1) % ckpt_every == 0 or it + 1 == steps):
            torch.save({"pos": pos.detach(), "opt": opt.state_dict(), "step": it + 1}, ck)
    rec["history"] = hist
    rec["step_s_mean"] = float(np.mean(step_times)) if step_times else None
    # the first step includes the first use of the backward's kernels
    rec["step_s_median"] = float(np.median(step_times)) if step_times else None
    rec["band_units_per_step"] = int(bidx.numel()) * A * cfg.n_freq
    rec["total_s"] = time.perf_counter() - t0
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, f"pipeline_config{cid}.json"), "w") as f:
            json.dump(rec, f, indent=1)
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--sigma", type=float, default=3e-4)
    ap.add_argument("--tau-vox", type=float, default=1.5)
    ap.add_argument("--lr", type=float, default=2e-6)
    ap.add_argument("--objective", choices=("energy", "sharpness"), default="energy")
    ap.add_argument("--K", type=int, default=64)
    ap.add_argument("--out", default=None)
    ap.add_argument("--ckpt-every", type=int, default=10)
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--tau-search", action="store_true",
                    help="run the isosurface offset search (ray-marched channels) after the BFS")
    ap.add_argument("--rays", type=int, default=512, help="rays per side per antenna")
    ap.add_argument("--fov", type=float, default=20.0, help="ray half-angle, degrees")
    ap.add_argument("--num-rx", type=int, default=50)
    ap.add_argument("--repeat-prep", type=int, default=0, metavar="N",
                    help="time the labelling, start voxels and BFS N more times (warm)")
    ap.add_argument("--variant", choices=("rec", "profile"), default="rec",
                    help="forward/backward pair: the recurrence or the range-profile pair")
    a = ap.parse_args(argv)
    rec = run(a.config, a.steps, a.sigma, a.tau_vox, a.lr, a.K, a.out, a.ckpt_every, a.resume,
              objective=a.objective, tau_search=a.tau_search, rays=a.rays, fov=a.fov,
              num_rx=a.num_rx, variant=a.variant, repeat_prep=a.repeat_prep)
    print(json.dumps({k: v for k, v in rec.items() if k != "history"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
