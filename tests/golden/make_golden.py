"""Writes tests/golden/*.npz: small input/output vectors for the hot path.

The reference ships no tests, golden vectors or fixtures, and its modules
cannot be imported here (pycuda, cc3d, skimage, open3d and trimesh are not
installed), so these vectors come from the pure-Python restatements in
oracle/oracle.py (py_bp_ref, py_sn_ref, py_bfs_fifo) -- line-by-line
restatements of imaging_gpu.cu:48-122, surface_normal_est.cu:64-149 and
sdf.py:180-222 -- plus an FP64 brute-force sum written with numpy
broadcasting.  They freeze those restatements so that later changes to the
C oracle or the kernels are checked against fixed numbers.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

C0 = 2.99792458e8


def bp_inputs(seed, n_tx=4, R=4, S=8, spectra=False):
    rng = np.random.default_rng(seed)
    lam = C0 / 77.5001e9
    rxo = np.array([[-0.005 - lam / 2 * k, 0.0, 0.0] for k in (3, 2, 1, 0)])[:R]
    tx = np.stack([rng.uniform(-0.02, 0.02, n_tx), rng.uniform(-0.02, 0.02, n_tx),
                   rng.uniform(0.25, 0.3, n_tx)], 1)
    ant = np.repeat(tx, R, axis=0)
    if spectra:
        S = 512
    meas = (rng.normal(size=(n_tx * R, S)) + 1j * rng.normal(size=(n_tx * R, S)))
    xl = np.linspace(-0.01, 0.01, 5)
    yl = np.linspace(-0.008, 0.008, 4)
    zl = np.linspace(-0.005, 0.005, 3)
    return xl, yl, zl, ant, meas.astype(np.complex64), rxo, lam


def brute_f64(xl, yl, zl, ant, meas, rxo, f0, df, is_ti):
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    X, Y, Z = np.meshgrid(f32(xl), f32(yl), f32(zl), indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)               # (V,3)
    p = f32(ant)                                                      # (A,3)
    o = f32(rxo)[np.arange(len(ant)) % len(rxo)]                      # (A,3)
    d = (np.linalg.norm(v[:, None] - p[None], axis=2)
         + np.linalg.norm(v[:, None] - p[None] - o[None], axis=2))    # (V,A)
    if is_ti:
        d = d + 0.15
    f = f0 + df * np.arange(meas.shape[1])
    ph = np.exp(-2j * np.pi * d[:, :, None] * f[None, None, :] / C0)  # (V,A,F)
    c = np.einsum("vaf,af->va", ph, meas.astype(np.complex128))
    vec = (p + o / 2)[None] - v[:, None]                               # (V,A,3)
    return c.sum(1), np.einsum("vak,va->vk", vec, c)


def main():
    out = {}
    slope = 5861160.0
    for name, seed, ti in (("a", 1, False), ("b", 2, True)):
        xl, yl, zl, ant, meas, rxo, lam = bp_inputs(seed)
        full = O.py_bp_ref(xl, yl, zl, ant, meas, rxo, slope, lam, is_ti=ti)
        f0 = C0 / float(np.float32(lam))
        img64, mom64 = brute_f64(xl, yl, zl, ant, meas, rxo, f0, float(np.float32(slope)), ti)
        out.update({f"{name}_xl": xl, f"{name}_yl": yl, f"{name}_zl": zl, f"{name}_ant": ant,
                    f"{name}_meas": meas, f"{name}_rxo": rxo, f"{name}_lam": lam,
                    f"{name}_ti": ti, f"{name}_full": full, f"{name}_f64": img64,
                    f"{name}_mom64": mom64})
        xl, yl, zl, ant, spec, rxo, lam = bp_inputs(seed + 10, spectra=True)
        fs = np.float32(3e8 / (2 * 3.0009e9) * 64 / 512)
        interp = O.py_bp_ref(xl, yl, zl, ant, spec, rxo, slope, lam, fs, is_ti=ti, interp=True)
        vidx = np.flatnonzero(np.abs(interp) > 0.2 * np.abs(interp).max())
        sn = O.py_sn_ref(vidx, interp, xl, yl, zl, ant, spec, rxo, lam, fs, is_ti=ti)
        out.update({f"{name}_spec": spec, f"{name}_fs": fs, f"{name}_interp": interp,
                    f"{name}_vidx": vidx, f"{name}_sn": sn, f"{name}_ixl": xl,
                    f"{name}_iyl": yl, f"{name}_izl": zl, f"{name}_iant": ant})
    np.savez_compressed(os.path.join(HERE, "bp_tiny.npz"), **out)

    rng = np.random.default_rng(5)
    shape = (7, 6, 5)
    n = rng.normal(size=shape + (3,))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    mask = rng.uniform(size=shape) < 0.6
    n[~mask] = np.nan
    lab = np.where(mask, 2, 1)
    lab[:, :, 3:][mask[:, :, 3:]] = 3  # two labels sharing the mask
    seeds = [int(np.flatnonzero(((lab == L) & mask).ravel())[0]) for L in (2, 3)]
    rsdf, hops = O.py_bfs_fifo(n, lab, seeds, 1e-3)
    np.savez_compressed(os.path.join(HERE, "bfs_tiny.npz"), normals=n, labels=lab,
                        seeds=np.array(seeds), dx=1e-3, rsdf=rsdf, hops=hops)
    print("wrote bp_tiny.npz, bfs_tiny.npz")


if __name__ == "__main__":
    main()
