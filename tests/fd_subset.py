"""FP64 central differences of L = sum_{v in S} w_v |I_v|^2 over a voxel
subset S, I = scale * sum_ij s_ij exp(-j 2 pi f_j d_iv / c) (monostatic
rows), with respect to single antenna coordinates, and the per-coordinate
bound that sets the tolerance of an FP32 gradient.

Moving antenna i by h along axis k changes only row i's contribution, so
L(p + h e_ik) comes from the FP64 image over S (the oracle) minus row i's
FP64 contribution at p plus its contribution at the moved position.  The
bound B_ik = sum_v 2 w_v |I_v| scale |dd_iv/dp_ik| sum_j k_j |s_ij| is the
sum of the magnitudes of the terms of dL/dp_ik: an evaluation whose every
term has relative error e has absolute error at most e * B_ik.
"""
import numpy as np

C0 = 2.99792458e8


def voxels_of(sc, vidx):
    ny, nz = len(sc.y_locs), len(sc.z_locs)
    ix, r = np.divmod(np.asarray(vidx, np.int64), ny * nz)
    iy, iz = np.divmod(r, nz)
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    return np.stack([f(sc.x_locs)[ix], f(sc.y_locs)[iy], f(sc.z_locs)[iz]], 1)


def row64(sc, v, p, i):
    """Row i's FP64 contribution sum_j s_ij exp(-j k_j d) at the voxels v from position p."""
    f = sc.f0 + sc.df * np.arange(sc.meas.shape[1])
    d = 2 * np.linalg.norm(v - p[None, :], axis=1)
    return np.exp(-2j * np.pi * d[:, None] * f[None, :] / C0) @ sc.meas[i].astype(np.complex128)


def fd_and_bound(oracle, sc, ant, vidx, w, scale, coords, h=1e-6):
    """(central differences, bounds) for each (row, axis) of coords; ant holds
    FP64 positions that are exact in FP32."""
    base = oracle.bp_f64(sc.x_locs, sc.y_locs, sc.z_locs, ant, sc.meas, np.zeros((1, 3)),
                         sc.f0, sc.df, vidx=vidx)
    v = voxels_of(sc, vidx)
    I = scale * base
    kj = 2 * np.pi * (sc.f0 + sc.df * np.arange(sc.meas.shape[1])) / C0
    fd, bound = [], []
    for (i, k) in coords:
        r0 = row64(sc, v, ant[i], i)
        vals = []
        for sgn in (1.0, -1.0):
            p = ant[i].copy()
            p[k] += sgn * h
            img = base - r0 + row64(sc, v, p, i)
            vals.append(float(np.sum(w * np.abs(scale * img) ** 2)))
        fd.append((vals[0] - vals[1]) / (2 * h))
        dd = 2 * np.abs(v[:, k] - ant[i, k]) / np.linalg.norm(v - ant[i][None, :], axis=1)
        bound.append(float(np.sum(2 * w * np.abs(I) * scale * dd) *
                           np.sum(kj * np.abs(sc.meas[i]))))
    return np.array(fd), np.array(bound)
