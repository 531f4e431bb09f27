"""Benchmark of the image-formation hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--K 256]

A step is one full-frequency backprojection of the configuration's synthetic
scene with the fused normal-vote moment (recurrence kernel, FP64 re-anchor
every --K samples: 256 by default, as 64-sample FP32 Horner chains combined
in r^64; the library's own default is 64), inputs resident in HBM.  Prints
ONE JSON line:

* value      W / t  in G voxel*row*sample/s (W = V*A*F, t = mean step time)
* roofline   FP32 VALU: algorithmic FLOPs V*A*(8F+14) per launch over the
             launch's mean duration from HIP events on the launch stream,
             against the MI355X FP32 vector peak (157.3 TFLOP/s: one wave64
             v_fma_f32 per SIMD every 2 cycles at 2.4 GHz), and against the
             119 TFLOP/s a Horner-shaped FMA stream sustains on MI355X with
             bounded data (power-limited to a 2.0 GHz clock; a constant, not
             measured by this run); traffic from the committed
             rocprofv3 --pmc summary when present (its source is named)
* accuracy   the timed output against the FP64 oracle on a seeded voxel
             sample that includes the strongest voxels (outside the timed
             region): relative L2 of image and moment, the normal angle
* rec_K64    the same step with the library's default re-anchor interval
             (K = 64) when --K differs
* plain      the direct FP32 kernels on the same inputs (fewer steps):
             plain_fast (phase in revolutions, hardware sine/cosine of its
             fraction) is the optimised direct baseline; plain (library
             sincosf of the ~1000 rad phase) pays the full argument
             reduction.  speedup_vs_direct states both ratios and whether
             the 20x target holds against the optimised one
* config5_forward  one full-grid forward with the moment at the largest
             configuration (100x100 aperture, 512 samples, 256x256x128 voxels)
* profile    the same sum from per-row range-profile tables (mmn_bp_profile):
             its time, its speed-up over the recurrence and its difference from
             the recurrence's image and moment, and its grouped-series
             backward (mmn_bp_profile_backward)
* backward   one backward (dL/dmeas, dL/dant) of L = sum w|I|^2
* rsdf_bfs   GPU 6-connected labelling and GPU BFS over the 5%-filtered vote
             field
* dropin_end_to_end_ms  image_cuda (full branch) with NumPy in and out, the
             host-device copies included
* cpu_baseline  the CPU restatement of the reference's imaging loop
             (imaging.cpp:60-139, OpenMP, the host threads this process may
             use) on a bounded voxel sample; .numpy_1core the NumPy
             restatement of imaging_gpu.cu:48-122 with the vote moment, FP64
             and FP32, on one core; .bfs the C restatement of the reference's
             FIFO BFS on one core and the Python deque restatement (the
             reference's own formulation), on the same field, and whether the
             GPU BFS equals them bit for bit
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12
# FMA-only microbenchmark on MI355X with the recurrence's own operand pattern
# (complex Horner steps of two voxels per lane, half the FMAs taking a sample
# from an SGPR, non-packed v_fma_f32, 8 waves per SIMD) and bounded data (the
# step on the unit circle): 117.9-119.2 TFLOP/s sustained over 120 ms, the
# shader clock power-limited to 1.99-2.00 GHz; a constant, not re-measured.
FMA_CEILING_MEASURED = 119.0e12
FMA_CEILING_SOURCE = ("Horner-shaped FMA microbenchmark on MI355X (two complex chains per "
                      "lane, half the FMAs with an SGPR operand, bounded data): 117.9-119.2 "
                      "TFLOP/s sustained at a power-limited 2.0 GHz, measured once, not by "
                      "this run")


def flops_fwd(V, A, F):
    return V * A * (8.0 * F + 14.0)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """(threads used, nproc, CPUs in this process's affinity mask, OMP_NUM_THREADS
    as given).  OpenMP runs on OMP_NUM_THREADS when the environment sets it
    (a GPU host may allot a share of its cores per GPU), else on every CPU of
    the affinity mask."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    env = os.environ.get("OMP_NUM_THREADS")
    threads = int(env) if env and env.isdigit() and int(env) > 0 else aff
    return threads, nproc, aff, env


def cpu_baseline(sc, target_s=12.0):
    """Time the CPU imaging restatement on a seeded voxel sample (cpu_baseline leg)."""
    from oracle import oracle as O
    cores, nproc, aff, env = host_threads()
    rng = np.random.default_rng(99)
    V = len(sc.x_locs) * len(sc.y_locs) * len(sc.z_locs)
    A, F = sc.meas.shape
    rxo = np.zeros((1, 3))
    n = cores * 2
    while True:
        vidx = np.sort(rng.choice(V, size=n, replace=False))
        t0 = time.perf_counter()
        O.bp_cpu_imaging(sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas, rxo, sc.slope,
                         sc.wavelength, vidx=vidx)
        dt = time.perf_counter() - t0
        if dt >= target_s or n >= V:
            break
        n = int(min(V, max(n * 2, n * target_s / max(dt, 1e-3) * 1.05)))
    units = n * A * F
    return {"value": units / dt / 1e9, "unit": "G voxel*row*sample/s", "cores": cores,
            "kind": "port",
            "sample": f"{n} random voxels of the {V}-voxel grid x {A} rows x {F} samples "
                      f"({dt:.1f} s, {cores} OpenMP threads, {cpu_model()})",
            "nproc": nproc, "affinity_cpus": aff, "omp_num_threads_env": env,
            "cpu_model": cpu_model()}


def numpy_baseline(sc, target_s=4.0):
    """CPU-1: the NumPy restatement of the full branch with the vote moment,
    FP64 and FP32, on one core, each on a seeded voxel sample sized to about
    target_s seconds."""
    from oracle import oracle as O
    V = len(sc.x_locs) * len(sc.y_locs) * len(sc.z_locs)
    A, F = sc.meas.shape
    out = {}
    for name, dt_ in (("fp64", np.float64), ("fp32", np.float32)):
        rng = np.random.default_rng(98)
        n = 4
        while True:
            vidx = np.sort(rng.choice(V, size=n, replace=False))
            t0 = time.perf_counter()
            O.np_bp_full(sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas, np.zeros((1, 3)),
                         sc.slope, sc.wavelength, vidx, dtype=dt_)
            dt = time.perf_counter() - t0
            if dt >= 0.5 * target_s or n >= 4096:
                break
            n = int(min(4096, max(2 * n, n * target_s / max(dt, 1e-3))))
        out[name] = {"value": n * A * F / dt / 1e9, "unit": "G voxel*row*sample/s",
                     "cores": 1, "kind": "port",
                     "sample": f"{n} random voxels x {A} rows x {F} samples ({dt:.1f} s), "
                               f"NumPy restatement of imaging_gpu.cu:48-122 with the vote "
                               f"moment, {name}"}
    return out


def accuracy_leg(sc, img, mom, n_rand=192, n_top=64):
    """The timed output against the FP64 oracle O64 on a seeded sample: n_rand
    random voxels and the n_top strongest.  Relative L2 of the image and the
    moment over the sample, the largest relative |I| error on the strongest,
    and the largest angle between the GPU's and O64's normals there."""
    import torch
    from oracle import oracle as O
    flat = img.reshape(-1)
    V = flat.numel()
    mag = flat.abs()
    top = torch.topk(mag, n_top).indices.cpu().numpy()
    rnd = np.random.default_rng(20260002).choice(V, size=n_rand, replace=False)
    vidx = np.unique(np.concatenate([top, rnd])).astype(np.int64)
    sel = torch.from_numpy(vidx).to(flat.device)
    g_img = flat[sel].cpu().numpy().astype(np.complex128)
    g_mom = mom.reshape(-1, 3)[sel].cpu().numpy().astype(np.complex128)
    t0 = time.perf_counter()
    w_img, w_mom = O.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas, np.zeros((1, 3)),
                            sc.f0, sc.df, vidx=vidx, moment=True)
    dt = time.perf_counter() - t0
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    is_top = np.isin(vidx, top)
    mag_err = np.abs(np.abs(g_img[is_top]) - np.abs(w_img[is_top])) / np.abs(w_img[is_top])

    def normals(i, m):
        n = (np.conj(i)[:, None] * m).real / np.abs(i)[:, None]
        return n / np.linalg.norm(n, axis=1, keepdims=True)
    cosang = np.clip(np.sum(normals(g_img[is_top], g_mom[is_top]) *
                            normals(w_img[is_top], w_mom[is_top]), 1), -1, 1)
    err = rel(g_img, w_img)
    return {"image_rel_l2": err, "moment_rel_l2": rel(g_mom, w_mom),
            "max_rel_mag_err_strongest": float(mag_err.max()),
            "max_normal_angle_deg_strongest": float(np.degrees(np.arccos(cosang)).max()),
            "meets_1e-4": bool(err <= 1e-4),
            "sample": f"{len(vidx)} voxels: the {n_top} strongest and {n_rand} seeded random; "
                      f"FP64 oracle O64 ({dt:.1f} s on the host)"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--K", type=int, default=256,
                    help="FP64 re-anchor interval of the recurrence (8-64: one FP32 Horner "
                         "chain; 128, 256: chains of 64 combined in r^64)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--profile-steps", type=int, default=5,
                    help="timed launches of the range-profile forward (0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="forward only")
    ap.add_argument("--bfs-only", type=int, default=0, metavar="N",
                    help="build the configuration's BFS field (range-profile forward with "
                         "the moment, 5%% filter, vote, labelling, start voxels), time N "
                         "calls of the RSDF BFS and print them as the JSON line")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="after the timed region, copy a BYTES-sized buffer once (a copy "
                         "kernel of known size for scaling rocprofv3 FETCH/WRITE_SIZE)")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build measures one GPU; the path shards by voxel "
                          "slabs across GPUs but no multi-GPU run is part of it",
                          "n_gpus": a.gpus}))
        return 2

    import torch
    from mmnorm_amd import ops, synth

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cfg = synth.CONFIGS[a.config]
    t0 = time.perf_counter()
    sc = synth.make_scene(a.config, device=dev)
    t_scene = time.perf_counter() - t0
    log(f"scene config {a.config} ready in {t_scene:.1f} s")
    T = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(dev)
    meas, ant = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = T(sc.x_locs, np.float32), T(sc.y_locs, np.float32), T(sc.z_locs, np.float32)
    V, A, F = cfg.nvox, cfg.rows, cfg.n_freq
    W = float(V) * A * F
    if a.bfs_only > 0:
        return bfs_only(sc, meas, ant, xl, yl, zl, dev, a.bfs_only)

    def fwd(variant):
        return ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, variant=variant,
                               K=a.K, moment=True)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t0 = time.perf_counter()
        for e0, e1 in evs:
            e0.record(st)
            out = fn()
            e1.record(st)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = [e0.elapsed_time(e1) for e0, e1 in evs]
        return wall, kms, out

    wall, kms, (img, mom) = timed(lambda: fwd("rec"), a.steps, a.warmup)
    ms_step = wall / a.steps * 1e3
    log(f"rec forward: {ms_step:.2f} ms/step")
    k_avg = float(np.mean(kms)) * 1e-3
    k_med = float(np.median(kms)) * 1e-3
    achieved = flops_fwd(V, A, F) / k_avg
    rec = {"metric": "G voxel*antenna*freq/s (recurrence fwd+vote, FP32 VALU)",
           "value": W / (wall / a.steps) / 1e9, "unit": "G voxel*row*sample/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (FP64 re-anchor)", "data": "synthetic (seeded scene, see synth.py)",
           "config": {"workload": f"config {cfg.cid}: {cfg.n_ap}x{cfg.n_ap} aperture, "
                                  f"{F} samples 77-81 GHz, {cfg.grid[0]}x{cfg.grid[1]}x"
                                  f"{cfg.grid[2]} voxels at {cfg.pitch * 1e3:g} mm, "
                                  f"{cfg.n_scat} scatterers",
                      "V": V, "A": A, "F": F, "K": a.K, "kernel": "k_bp_rec",
                      "parallelism": "single GPU"},
           "kernel_ms": k_avg * 1e3, "kernel_ms_median": k_med * 1e3,
           "roofline": {"bound": "valu", "achieved": achieved / 1e12,
                        "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32,
                        "frac_of_measured_fma_ceiling": achieved / FMA_CEILING_MEASURED,
                        "fma_ceiling_source": FMA_CEILING_SOURCE,
                        "traffic": None}}
    prof = os.path.join(ROOT, "profiles", f"traffic_bp_rec_config{cfg.cid}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            tr = json.load(f)
        rec["roofline"]["traffic"] = tr.get("bytes_per_launch")
        rec["roofline"]["traffic_source"] = (
            f"profiles/traffic_bp_rec_config{cfg.cid}.json: rocprofv3 --pmc FETCH_SIZE / "
            f"WRITE_SIZE of {tr.get('kernel')} at tree {tr.get('tree', 'unrecorded')}, "
            f"not measured by this run")
    log("accuracy leg (FP64 oracle on a voxel sample) ...")
    rec["accuracy"] = accuracy_leg(sc, img, mom)
    log(f"accuracy: {rec['accuracy']}")

    if not a.no_extras:
        if a.K != 64:
            # the same step with the library's default re-anchor interval
            _, k64, _ = timed(lambda: ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0,
                                                      df=sc.df, K=64, moment=True), 3, 1)
            k64_s = float(np.mean(k64)) * 1e-3
            rec["rec_K64"] = {"value": W / k64_s / 1e9, "kernel_ms": k64_s * 1e3,
                              "roofline_frac": flops_fwd(V, A, F) / k64_s / PEAK_FP32}
            log(f"rec forward, K = 64: {k64_s * 1e3:.2f} ms")
        if a.plain_steps > 0:
            pw, pk, _ = timed(lambda: fwd("plain"), a.plain_steps, 1)
            p_s = float(np.mean(pk)) * 1e-3
            log(f"plain forward: {p_s * 1e3:.2f} ms")
            rec["plain"] = {"value": W / p_s / 1e9, "kernel_ms": p_s * 1e3,
                            "tflops": flops_fwd(V, A, F) / p_s / 1e12,
                            "rec_over_plain": p_s / k_avg,
                            "kind": "library sincosf of the FP32 phase (radians)"}
            fw, fk, _ = timed(lambda: fwd("plain_fast"), max(a.plain_steps, 2), 1)
            f_s = float(np.mean(fk)) * 1e-3
            log(f"plain_fast forward: {f_s * 1e3:.2f} ms")
            rec["plain_fast"] = {"value": W / f_s / 1e9, "kernel_ms": f_s * 1e3,
                                 "rec_over_plain_fast": f_s / k_avg,
                                 "kind": "hardware sin/cos of the FP32 phase fraction "
                                         "(revolutions)"}
        if a.profile_steps > 0:
            bb = ops.grid_bbox(xl, yl, zl)
            pw, pk, (pimg, pmom) = timed(
                lambda: ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df,
                                        variant="profile", moment=True, bbox=bb),
                a.profile_steps, 1)
            q_s = float(np.mean(pk)) * 1e-3
            diff = lambda u, v: float(torch.linalg.norm(u - v) / torch.linalg.norm(v))
            rec["profile"] = {"value": W / q_s / 1e9, "kernel_ms": q_s * 1e3,
                              "speedup_over_rec": k_avg / q_s,
                              "rel_diff_vs_rec": diff(pimg, img), "moment_rel_diff_vs_rec":
                              diff(pmom, mom),
                              "kind": "same sum from per-row cubic range-profile tables "
                                      "(mmn_bp_profile); value = voxel*row*sample / time"}
            log(f"profile forward: {q_s * 1e3:.2f} ms, {k_avg / q_s:.1f}x the recurrence")
            # its backward for L = sum w |I|^2 (I = img / A), both outputs
            wq = torch.rand(V, device=dev, generator=torch.Generator(dev).manual_seed(1))
            gq = (2.0 * wq * pimg.reshape(-1) / A / A).to(torch.complex64).contiguous()
            qw, qk, _ = timed(lambda: ops.backproject_backward(
                gq, meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, scale=1.0 / A,
                variant="profile", bbox=bb), a.profile_steps, 1)
            qb = float(np.mean(qk)) * 1e-3
            rec["profile"].update({"backward_ms": qb * 1e3, "backward_over_forward": qb / q_s})
            log(f"profile backward: {qb * 1e3:.2f} ms")
        if a.plain_steps > 0:
            sp = {"rec_over_plain_fast": f_s / k_avg, "rec_over_plain_sincosf": p_s / k_avg,
                  "rec_K64_over_plain_fast": (f_s / k64_s) if a.K != 64 else f_s / k_avg,
                  "target": 20.0, "target_met_vs_plain_fast": bool(f_s / k_avg >= 20.0),
                  "target_met_vs_plain_sincosf": bool(p_s / k_avg >= 20.0),
                  "denominator": "plain_fast: the optimised direct FP32 kernel (hardware "
                                 "sin/cos of the phase fraction; the same accuracy as "
                                 "plain, whose library sincosf of the ~1000 rad phase pays "
                                 "the full argument reduction)"}
            if a.profile_steps > 0:
                sp["profile_over_plain_fast"] = f_s / q_s
                # the same sum from range-profile tables that the recurrence
                # builds (7e-7 from the recurrence's image): not the metric's
                # kernel, reported beside it
                sp["target_met_by_profile_vs_plain_fast"] = bool(f_s / q_s >= 20.0)
            rec["speedup_vs_direct"] = sp
        rec["config5_forward"] = config5_stage(dev, a.K)
        log(f"config 5 forward: {rec['config5_forward']}")
        rec["config4_backward"] = config4_stage(dev)
        log(f"config 4 forward/backward: {rec['config4_backward']}")
        rec["interpolated"] = interp_stage(sc, ant, xl, yl, zl, dev)
        log(f"interpolated branch: {rec['interpolated']}")
        # backward of L = sum w |I|^2 (I = img / A)
        w = torch.rand(V, device=dev, generator=torch.Generator(dev).manual_seed(1))
        g = (2.0 * w * img.reshape(-1) / A / A).to(torch.complex64).contiguous()
        bw, bk, _ = timed(lambda: ops.backproject_backward(g, meas, ant, xl, yl, zl, f0=sc.f0,
                                                           df=sc.df, scale=1.0 / A, K=32),
                          2, 1)
        b_s = float(np.mean(bk)) * 1e-3
        rec["backward"] = {"ms": b_s * 1e3, "over_forward": b_s / k_avg,
                           "tflops": V * A * (16.0 * F + 30.0) / b_s / 1e12}
        # the position gradient alone (what autograd asks for when the
        # measurements take no gradient, as in the pipeline's steps)
        _, pk2, _ = timed(lambda: ops.backproject_backward(
            g, meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, scale=1.0 / A,
            need_meas=False), 2, 1)
        rec["backward"]["ant_only_ms"] = float(np.mean(pk2))
        log(f"backward: {b_s * 1e3:.2f} ms (position gradient alone "
            f"{rec['backward']['ant_only_ms']:.2f} ms)")
        rec["rsdf_bfs"], bfs_state = bfs_stage(img, mom, sc, dev)
        log(f"rsdf bfs: {rec['rsdf_bfs']}")
        # the drop-in end to end: NumPy in, NumPy out (host-device copies included)
        from mmnorm_amd import py_image_gpu
        lam = synth.C0 / sc.f0
        dropin = lambda: py_image_gpu.image_cuda(
            sc.x_locs, sc.y_locs, sc.z_locs, sc.ant, sc.meas, np.zeros((1, 3)), sc.df, lam,
            sc.df * (F - 1), F, use_interpolated_processing=False, return_phase_centers=False)
        dropin()
        t0 = time.perf_counter()
        for _ in range(2):
            dropin()
        rec["dropin_end_to_end_ms"] = (time.perf_counter() - t0) / 2 * 1e3
        log(f"image_cuda drop-in end to end: {rec['dropin_end_to_end_ms']:.1f} ms")
    rec["scene_s"] = t_scene
    if a.copy_calibrate > 0:
        src = torch.empty(a.copy_calibrate // 4, dtype=torch.float32, device=dev).fill_(1.0)
        torch.cuda.synchronize()
        dst = src.clone()
        torch.cuda.synchronize()
        rec["copy_calibrate_bytes"] = int(dst.numel() * 4)
        log(f"copy calibration: {dst.numel() * 4} bytes copied")
    if not a.no_cpu:
        log("cpu baseline ...")
        rec["cpu_baseline"] = cpu_baseline(sc)
        log("numpy cpu-1 baseline ...")
        rec["cpu_baseline"]["numpy_1core"] = numpy_baseline(sc)
        if not a.no_extras:
            rec["cpu_baseline"]["bfs"] = cpu_bfs_baseline(bfs_state)
            log(f"cpu bfs: {rec['cpu_baseline']['bfs']}")
    print(json.dumps(rec), flush=True)
    return 0


def config5_stage(dev, K):
    """The largest configuration (100x100 aperture, 512 samples, 256x256x128
    voxels at 1 mm, 20,000 scatterers): one full-grid recurrence forward with
    the vote moment after a warm-up launch, and the range-profile forward,
    HIP-event timed."""
    import torch
    from mmnorm_amd import ops, synth
    t0 = time.perf_counter()
    sc = synth.make_scene(5, device=dev)
    scene_s = time.perf_counter() - t0
    T = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(dev)
    meas, ant = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    st = torch.cuda.current_stream()

    def ev_ms(f):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        out = f()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), out
    fwd = lambda: ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, K=K, moment=True)
    # warm-up on a voxel list (same kernel, a fraction of the work)
    ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, K=K, moment=True,
                    vox_idx=torch.arange(4096, dtype=torch.int32, device=dev))
    ms, (img, _) = ev_ms(fwd)
    W = float(sc.cfg.work)
    bb = ops.grid_bbox(xl, yl, zl)
    prof = lambda: ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df,
                                   variant="profile", moment=True, bbox=bb)
    prof()
    pms, (pimg, _) = ev_ms(prof)
    diff = float(torch.linalg.norm(pimg - img) / torch.linalg.norm(img))
    return {"forward_ms": ms, "value": W / (ms * 1e-3) / 1e9, "unit": "G voxel*row*sample/s",
            "tflops": flops_fwd(sc.cfg.nvox, sc.cfg.rows, sc.cfg.n_freq) / (ms * 1e-3) / 1e12,
            "profile_forward_ms": pms, "profile_rel_diff_vs_rec": diff, "units": int(W),
            "scene_s": scene_s}


def config4_stage(dev):
    """The differentiable configuration at full size (40x40 aperture, 128
    samples, 96^3 voxels): forward (image only) and backward (dL/dmeas and
    dL/dant of L = sum w |I|^2) times."""
    import torch
    from mmnorm_amd import ops, synth
    sc = synth.make_scene(4, device=dev)
    T = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x, dtype=dt)).to(dev)
    meas, ant = T(sc.meas, np.complex64), T(sc.ant, np.float32)
    xl, yl, zl = (T(v, np.float32) for v in (sc.x_locs, sc.y_locs, sc.z_locs))
    A = ant.shape[0]

    def t_of(f, n=3):
        out = f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            out = f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3, out
    f_ms, img = t_of(lambda: ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, K=64))
    w = torch.rand(img.numel(), device=dev, generator=torch.Generator(dev).manual_seed(2))
    g = (2.0 * w * img.reshape(-1) / A / A).to(torch.complex64).contiguous()
    b_ms, _ = t_of(lambda: ops.backproject_backward(g, meas, ant, xl, yl, zl, f0=sc.f0,
                                                    df=sc.df, scale=1.0 / A, K=32))
    bb = ops.grid_bbox(xl, yl, zl)
    pf_ms, _ = t_of(lambda: ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df,
                                            variant="profile", bbox=bb))
    pb_ms, _ = t_of(lambda: ops.backproject_backward(g, meas, ant, xl, yl, zl, f0=sc.f0,
                                                     df=sc.df, scale=1.0 / A,
                                                     variant="profile", bbox=bb))
    return {"forward_ms": f_ms, "backward_ms": b_ms, "over_forward": b_ms / f_ms,
            "profile_forward_ms": pf_ms, "profile_backward_ms": pb_ms,
            "profile_over_forward": pb_ms / pf_ms, "units": int(sc.cfg.work)}


def interp_stage(sc, ant, xl, yl, zl, dev):
    """The reference's default interpolated branch on the same grid and rows
    (4096 spectrum bins per row, seeded random spectra): the reference-arithmetic
    kernel (FP64 cos/sin per voxel and row) against the MI355X kernel, both with
    the image only, and the MI355X kernel with the vote moment fused."""
    import torch
    from mmnorm_amd import ops
    S = 4096
    g = torch.Generator(dev).manual_seed(5)
    spec = torch.randn(ant.shape[0], S, dtype=torch.complex64, device=dev, generator=g)
    lam = float(np.float32(3e8 / 77e9))
    fs = float(np.float32(3e8 / (2 * 4e9) * 512 / S))

    def t_of(f, n=3):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3
    ref_ms = t_of(lambda: ops.cuda_image(spec, ant, xl, yl, zl, None, slope=0.0, wavelength=lam,
                                         fft_spacing=fs, interpolated=True))
    fast_ms = t_of(lambda: ops.backproject_interp(spec, ant, xl, yl, zl, wavelength=lam,
                                                  fft_spacing=fs))
    mom_ms = t_of(lambda: ops.backproject_interp(spec, ant, xl, yl, zl, wavelength=lam,
                                                 fft_spacing=fs, moment=True))
    a = ops.cuda_image(spec, ant, xl, yl, zl, None, slope=0.0, wavelength=lam, fft_spacing=fs,
                       interpolated=True)
    b = ops.backproject_interp(spec, ant, xl, yl, zl, wavelength=lam, fft_spacing=fs)
    diff = float((torch.linalg.norm(a - b) / torch.linalg.norm(a)).item())
    return {"ref_ms": ref_ms, "fast_ms": fast_ms, "fast_with_moment_ms": mom_ms,
            "speedup": ref_ms / fast_ms, "rel_l2_fast_vs_ref": diff, "bins": S}


def bfs_only(sc, meas, ant, xl, yl, zl, dev, n):
    """Time n calls of the RSDF BFS on the configuration's field."""
    import torch
    from mmnorm_amd import ops
    img, mom = ops.backproject(meas, ant, xl, yl, zl, f0=sc.f0, df=sc.df, moment=True,
                               variant="profile", bbox=ops.grid_bbox(xl, yl, zl))
    normals, lab_t, valid, seeds, dx, _ = bfs_field(img, mom, sc, dev)
    seeds_t = torch.from_numpy(np.asarray(seeds, np.int32)).to(dev)
    ms = []
    for _ in range(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rsdf, hops = ops.rsdf_bfs(normals, lab_t, seeds_t, dx)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    rest = ms[1:] if n > 1 else ms
    print(json.dumps({"metric": "RSDF BFS wall time per call", "value": float(np.median(rest)),
                      "unit": "ms", "higher_is_better": False, "calls": n,
                      "first_call_ms": ms[0], "min_ms": float(min(rest)),
                      "voxels_reached": int((hops >= 0).sum().item()),
                      "levels": int(hops.max().item()) + 1,
                      "config": {"workload": f"config {sc.cfg.cid} field"}}), flush=True)
    return 0


def bfs_field(img, mom, sc, dev):
    """The BFS's inputs from an image and its vote moment: unit normals on the
    5 % filter, GPU 6-connected labels, the size rule and the start voxels."""
    import torch
    from mmnorm_amd import ops, sdf
    mag = img.abs().reshape(-1)
    vidx = torch.nonzero(mag > 0.05 * mag.max()).reshape(-1).to(torch.int32)
    m = mom.reshape(-1, 3)[vidx.long()].contiguous()
    nrm = ops.sn_from_moment(img.reshape(-1), m, vidx).double()
    nrm = nrm / torch.linalg.norm(nrm, dim=1, keepdim=True)
    shape = img.shape
    normals = torch.full((img.numel(), 3), float("nan"), dtype=torch.float64, device=dev)
    normals[vidx.long()] = nrm
    normals = normals.reshape(*shape, 3)
    dx = sc.x_locs[1] - sc.x_locs[0]
    sdf.find_components_gpu(normals)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    lab_t, areas = sdf.find_components_gpu(normals)   # GPU 6-connected labels + 1
    torch.cuda.synchronize()
    ccl_ms = (time.perf_counter() - tc) * 1e3
    valid = sdf.size_rule(areas, dx)
    seeds = sdf.component_seeds_gpu(normals, lab_t, valid, sc.x_locs, sc.y_locs, sc.z_locs)
    return normals, lab_t, valid, seeds, dx, (vidx, ccl_ms)


def bfs_stage(img, mom, sc, dev):
    """5% filter -> votes -> 6-connected components and the size rule (GPU) ->
    GPU FIFO BFS, timed."""
    import torch
    from mmnorm_amd import ops, sdf
    normals, lab_t, valid, _, dx, (vidx, ccl_ms) = bfs_field(img, mom, sc, dev)
    labeled = lab_t.cpu().numpy()
    nrm_np = normals.cpu().numpy()
    m1, m2, m3 = np.meshgrid(sc.y_locs, sc.x_locs, sc.z_locs)
    coords = np.concatenate((m2[..., None], m1[..., None], m3[..., None]), axis=-1)
    t0 = time.perf_counter()
    seeds = sdf.component_seeds(nrm_np, coords, labeled, valid)
    seeds_host_ms = (time.perf_counter() - t0) * 1e3
    # the device-side search on the same fields (a warm-up call, then timed)
    sdf.component_seeds_gpu(normals, lab_t, valid, sc.x_locs, sc.y_locs, sc.z_locs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seeds_dev = sdf.component_seeds_gpu(normals, lab_t, valid, sc.x_locs, sc.y_locs, sc.z_locs)
    seeds_ms = (time.perf_counter() - t0) * 1e3
    seeds_t = torch.from_numpy(seeds.astype(np.int32)).to(dev)
    ops.rsdf_bfs(normals, lab_t, seeds_t, dx)
    torch.cuda.synchronize()
    calls = []
    for _ in range(7):   # a call takes about a millisecond: the median of seven
        t0 = time.perf_counter()
        rsdf, hops = ops.rsdf_bfs(normals, lab_t, seeds_t, dx)
        torch.cuda.synchronize()
        calls.append(time.perf_counter() - t0)
    dt = float(np.median(calls))
    res = {"ms": dt * 1e3, "ms_min": min(calls) * 1e3, "calls": len(calls),
           "ccl_ms": ccl_ms, "seeds_ms": seeds_ms,
           "seeds_host_ms": seeds_host_ms,
           "seeds_equal_host": bool(np.array_equal(seeds_dev, seeds)),
           "voxels_reached": int((hops >= 0).sum().item()),
           "components": int(len(seeds)), "levels": int(hops.max().item()) + 1,
           "valid_voxels": int(vidx.numel())}
    # what the CPU baseline of the BFS needs: the same inputs and the GPU result
    state = (nrm_np, labeled, seeds, dx, rsdf.cpu().numpy(), hops.cpu().numpy())
    return res, state


def cpu_bfs_baseline(state):
    """The reference's FIFO BFS (sdf.py:180-222) restated in C, one core, on
    the same normals, labels and seeds; and whether the GPU result equals it
    bit for bit."""
    from oracle import oracle as O
    nrm_np, labeled, seeds, dx, rsdf_gpu, hops_gpu = state
    t0 = time.perf_counter()
    rsdf, hops = O.bfs_fifo(nrm_np, labeled, seeds, dx)
    dt = time.perf_counter() - t0
    same = bool(np.array_equal(hops, hops_gpu) and
                np.array_equal(rsdf.view(np.int64), rsdf_gpu.view(np.int64)))
    out = {"ms": dt * 1e3, "cores": 1, "kind": "port",
           "sample": f"the whole BFS of the bench's field ({int((hops >= 0).sum())} voxels "
                     f"reached), C restatement of sdf.py:180-222, {cpu_model()}",
           "gpu_bit_identical": same}
    # the reference's own formulation (deque + np.dot per voxel), when bounded
    if int((hops >= 0).sum()) <= 500_000:
        t0 = time.perf_counter()
        prs, phops = O.py_bfs_fifo(nrm_np, labeled, seeds, dx)
        out["python_ms"] = (time.perf_counter() - t0) * 1e3
        out["python_bit_identical"] = bool(
            np.array_equal(phops, hops_gpu) and
            np.array_equal(np.nan_to_num(prs, nan=-1.0), np.nan_to_num(rsdf_gpu, nan=-1.0)))
    return out


if __name__ == "__main__":
    sys.exit(main())
