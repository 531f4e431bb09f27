"""Seeded synthetic scenes for the five benchmark configurations.

Monostatic planar apertures (one RX, zero offset, no TI correction) centred
0.3 m above the scene centre; frequencies f_j = f0 + j df, linearly spaced
77-81 GHz; voxel grids centred on the origin, x_i = (i - (n-1)/2) * pitch.
Measurements follow the reference's sign convention (the correlator uses
exp(-j 2 pi f d / c), imaging_gpu.cu:112), so a point scatterer rho at x
contributes  rho * exp(+j 2 pi f * 2|x - a| / c)  to row a, sample f.  Noise
is circular complex Gaussian with power mean|s|^2 / 10^(SNR/10).

Random draws use numpy.random.default_rng(20260000 + config) in a fixed
order: scatterer positions, normals/reflectivities, noise.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

C0 = 2.99792458e8


@dataclasses.dataclass
class Config:
    cid: int
    n_ap: int            # aperture is n_ap x n_ap
    ap_pitch: float
    n_freq: int
    grid: tuple          # (NX, NY, NZ)
    pitch: float
    snr_db: float
    scene: str           # "points" | "ellipsoid_box" | "multi"
    n_scat: int
    height: float = 0.3
    f_lo: float = 77e9
    f_hi: float = 81e9

    @property
    def f0(self):
        return self.f_lo

    @property
    def df(self):
        return (self.f_hi - self.f_lo) / (self.n_freq - 1)

    @property
    def rows(self):
        return self.n_ap * self.n_ap

    @property
    def nvox(self):
        return int(np.prod(self.grid))

    @property
    def work(self):
        """voxel x row x sample units."""
        return self.nvox * self.rows * self.n_freq


CONFIGS = {
    1: Config(1, 20, 2e-3, 64, (64, 64, 32), 2e-3, 20.0, "points", 5),
    2: Config(2, 60, 2e-3, 256, (128, 128, 96), 1.5e-3, 20.0, "ellipsoid_box", 4000),
    3: Config(3, 20, 2e-3, 64, (64, 64, 32), 2e-3, 20.0, "points", 5),
    4: Config(4, 40, 2e-3, 128, (96, 96, 96), 2e-3, 20.0, "ellipsoid_box", 4000),
    5: Config(5, 100, 2e-3, 512, (256, 256, 128), 1e-3, 15.0, "multi", 20000),
}

# Shapes of the config-2/4 scene (metres): ellipsoid semi-axes and centre,
# box size and centre.
ELLIPSOID = ((0.04, 0.03, 0.02), (-0.03, 0.0, -0.01))
BOX = ((0.04, 0.04, 0.03), (0.045, 0.0, -0.015))


def axis(n, pitch):
    return (np.arange(n) - (n - 1) / 2.0) * pitch


def aperture(n, pitch, height):
    a = axis(n, pitch)
    gx, gy = np.meshgrid(a, a, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), np.full(n * n, height)], axis=1)


def _ellipsoid_surface(rng, n, semi, centre):
    """Area-uniform samples on an ellipsoid (rejection on the area element)."""
    a, b, c = semi
    pts, nrm = [], []
    bound = max(b * c, a * c, a * b)
    while sum(len(p) for p in pts) < n:
        u = rng.normal(size=(4 * n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        el = np.sqrt((b * c * u[:, 0]) ** 2 + (a * c * u[:, 1]) ** 2 + (a * b * u[:, 2]) ** 2)
        keep = rng.uniform(size=len(u)) * bound < el
        u = u[keep]
        p = u * np.array([a, b, c])
        g = p / np.array([a * a, b * b, c * c])
        pts.append(p + np.asarray(centre))
        nrm.append(g / np.linalg.norm(g, axis=1, keepdims=True))
    return np.concatenate(pts)[:n], np.concatenate(nrm)[:n]


def _box_surface(rng, n, size, centre):
    """Area-uniform samples on the six faces of an axis-aligned box."""
    sx, sy, sz = size
    areas = np.array([sy * sz, sy * sz, sx * sz, sx * sz, sx * sy, sx * sy])
    face = rng.choice(6, size=n, p=areas / areas.sum())
    uv = rng.uniform(-0.5, 0.5, size=(n, 3)) * np.array(size)
    ax = face // 2
    sign = np.where(face % 2 == 0, -1.0, 1.0)
    pts = uv.copy()
    pts[np.arange(n), ax] = sign * np.array(size)[ax] / 2
    nrm = np.zeros((n, 3))
    nrm[np.arange(n), ax] = sign
    return pts + np.asarray(centre), nrm


def _ellipsoid_area(semi):
    a, b, c = semi
    p = 1.6075  # Knud Thomsen's approximation
    return 4 * math.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def _box_area(size):
    sx, sy, sz = size
    return 2 * (sx * sy + sy * sz + sx * sz)


def scatterers(cfg: Config, rng):
    """Scatterer positions (n,3), reflectivities (n,) complex128, normals (n,3)."""
    look_from = np.array([0.0, 0.0, cfg.height])
    if cfg.scene == "points":
        ext = np.array([0.128, 0.128, 0.064])
        pos = rng.uniform(-0.5, 0.5, size=(cfg.n_scat, 3)) * ext
        rho = (rng.normal(size=cfg.n_scat) + 1j * rng.normal(size=cfg.n_scat)) / math.sqrt(2)
        return pos, rho, np.tile([0.0, 0.0, 1.0], (cfg.n_scat, 1))
    if cfg.scene == "ellipsoid_box":
        shapes = [("e",) + ELLIPSOID, ("b",) + BOX]
    else:  # three random ellipsoids and two random boxes inside the grid
        half = np.array(cfg.grid) * cfg.pitch / 2
        shapes = []
        for k in range(5):
            if k < 3:
                semi = rng.uniform([0.02, 0.02, 0.012], [0.045, 0.045, 0.03])
                ctr = rng.uniform(-half + semi + 0.005, half - semi - 0.005)
                shapes.append(("e", tuple(semi), tuple(ctr)))
            else:
                size = rng.uniform([0.03, 0.03, 0.02], [0.06, 0.06, 0.04])
                ctr = rng.uniform(-half + size / 2 + 0.005, half - size / 2 - 0.005)
                shapes.append(("b", tuple(size), tuple(ctr)))
    areas = np.array([_ellipsoid_area(s[1]) if s[0] == "e" else _box_area(s[1]) for s in shapes])
    counts = np.floor(cfg.n_scat * areas / areas.sum()).astype(int)
    counts[0] += cfg.n_scat - counts.sum()
    pos, nrm = [], []
    for (kind, dims, ctr), n in zip(shapes, counts):
        p, q = (_ellipsoid_surface if kind == "e" else _box_surface)(rng, int(n), dims, ctr)
        pos.append(p)
        nrm.append(q)
    pos, nrm = np.concatenate(pos), np.concatenate(nrm)
    look = look_from - pos
    look /= np.linalg.norm(look, axis=1, keepdims=True)
    cosang = np.maximum(np.sum(nrm * look, axis=1), 0.0)
    n = len(pos)
    cn = (rng.normal(size=n) + 1j * rng.normal(size=n)) / math.sqrt(2)
    return pos, cosang * cn, nrm


def synthesize(ant, pos, rho, freqs, device=None, chunk_rows=None):
    """s[a, f] = sum_p rho_p exp(+j 2 pi f 2|x_p - a| / c), complex128.

    Uses torch float64 on ``device`` when given (large scenes), NumPy otherwise.
    """
    k = 2 * math.pi * np.asarray(freqs, np.float64) / C0
    if device is None:
        out = np.zeros((len(ant), len(k)), np.complex128)
        step = chunk_rows or max(1, int(2e7 // max(1, len(pos) * len(k))))
        for lo in range(0, len(ant), step):
            a = ant[lo:lo + step]
            d = 2 * np.linalg.norm(a[:, None, :] - pos[None, :, :], axis=2)  # (ra, p)
            ph = np.exp(1j * d[:, :, None] * k[None, None, :])              # (ra, p, f)
            out[lo:lo + step] = np.einsum("p,rpf->rf", rho, ph)
        return out
    import torch
    t_ant = torch.as_tensor(ant, dtype=torch.float64, device=device)
    t_pos = torch.as_tensor(pos, dtype=torch.float64, device=device)
    t_rho = torch.as_tensor(rho, dtype=torch.complex128, device=device)
    t_k = torch.as_tensor(k, dtype=torch.float64, device=device)
    out = torch.zeros((len(ant), len(k)), dtype=torch.complex128, device=device)
    step = chunk_rows or max(1, int(1.5e8 // max(1, len(pos) * len(k))))
    for lo in range(0, len(ant), step):
        d = 2 * torch.cdist(t_ant[lo:lo + step], t_pos)                    # (ra, p)
        ph = d[:, :, None] * t_k[None, None, :]
        z = torch.polar(torch.ones_like(ph), ph)                             # exp(+j ph)
        out[lo:lo + step] = torch.einsum("p,rpf->rf", t_rho, z)
    return out.cpu().numpy()


@dataclasses.dataclass
class Scene:
    cfg: Config
    ant: np.ndarray        # (rows, 3) float64
    meas: np.ndarray       # (rows, F) complex64
    x_locs: np.ndarray     # float64
    y_locs: np.ndarray
    z_locs: np.ndarray
    pos: np.ndarray
    rho: np.ndarray
    normals: np.ndarray

    @property
    def f0(self):
        return self.cfg.f0

    @property
    def df(self):
        return self.cfg.df

    @property
    def wavelength(self):
        return C0 / self.cfg.f0

    @property
    def slope(self):
        return self.cfg.df


def make_scene(cid, device=None, cfg=None, seed=None) -> Scene:
    cfg = cfg or CONFIGS[cid]
    rng = np.random.default_rng(20260000 + cid if seed is None else seed)
    pos, rho, nrm = scatterers(cfg, rng)
    ant = aperture(cfg.n_ap, cfg.ap_pitch, cfg.height)
    freqs = cfg.f0 + cfg.df * np.arange(cfg.n_freq)
    s = synthesize(ant, pos, rho, freqs, device=device)
    p_sig = float(np.mean(np.abs(s) ** 2))
    sigma = math.sqrt(p_sig / 10 ** (cfg.snr_db / 10) / 2)
    s = s + sigma * (rng.normal(size=s.shape) + 1j * rng.normal(size=s.shape))
    nx, ny, nz = cfg.grid
    return Scene(cfg, ant, s.astype(np.complex64), axis(nx, cfg.pitch), axis(ny, cfg.pitch),
                 axis(nz, cfg.pitch), pos, rho, nrm)
