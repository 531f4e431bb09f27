"""Closed-form cases of the interpolated (range-bin) branch and its vote.

Each case is built so that its expected value follows from the geometry, the
FFT's definition and the reference's stated rules, not from the kernels'
expressions: imaging_gpu.cu:90-105 (bin floor(d / fs / 2), skip d > fs * S,
phasor exp(-j 2 pi d / lambda)), surface_normal_est.cu:110-149 (weight
Re(c conj I) / |I| on the vector from the voxel to p + o / 2) and
image_processor.py:22-35 (zero-padded FFT of the chirp samples) with
fs = 3e8 / (2 B) * N_samples / N_fft (py_image_gpu.py:270).  Used by
tests/test_oracle_interp.py (the CPU oracle) and tests/test_gpu_interp_closed.py
(the HIP kernels).
"""
from __future__ import annotations

import numpy as np

C0 = 2.99792458e8
H = 0.5                   # antenna height of the bin cases
FS = 2.0 ** -7            # a power-of-two fft_spacing: d / fs / 2 is exact in FP32
LAM = float(np.float32(C0 / 77e9))


def bin_case(S=64, is_ti=False):
    """One row at (0, 0, H), monostatic, voxels straight below it at
    z = H - m 2^-20 (m > 0 integer; exact in FP32).  The FP32 round trip is
    exactly 2 m 2^-20: the differences are exact, sqrt of the rounded square
    of a float returns that float (binary arithmetic, round to nearest), and
    o = 0 makes both legs the same.  So the reference's bin is
    floor((2 m 2^-20) / 2^-7 / 2) = m // 2^13 exactly, and the row is skipped
    iff d > fs S, i.e. m > S 2^12.  With is_ti the distance is d + 0.15
    rounded to FP32 (a float += 0.15 in double), and the bin is that times 64,
    floored (fs a power of two: the scaling is exact).  Returns the axes,
    row, spectrum, distances, bins, kept flags and the expected raw image."""
    ms = []
    for b0 in (0, 1, 7, 31):
        ms += [8192 * b0 - 1, 8192 * b0, 8192 * b0 + 1, 8192 * b0 + 4096, 8192 * (b0 + 1) - 1]
    edge = S * 4096                       # d == fs * S exactly: kept, bin S / 2
    ms += [edge - 1, edge, edge + 1, edge + 8192]
    ms = np.array(sorted(m for m in set(ms) if m > 0), np.int64)
    z64 = H - ms.astype(np.float64) * 2.0 ** -20
    z = z64.astype(np.float32)
    assert np.all(z.astype(np.float64) == z64)
    d = 2.0 * ms.astype(np.float64) * 2.0 ** -20
    if is_ti:
        d = (d + 0.15).astype(np.float32).astype(np.float64)
        scaled = d * 64.0                 # d / fs / 2, exact
        bins = np.floor(scaled).astype(np.int64)
        keep = d <= FS * S
    else:
        bins = ms // 8192
        keep = ms <= S * 4096
    j = np.arange(S)
    spec = ((j + 1) * (1 + 0.5j) * np.exp(1j * 0.37 * j)).astype(np.complex64)
    want = np.zeros(len(ms), np.complex128)
    sp64 = spec.astype(np.complex128)
    want[keep] = sp64[bins[keep]] * np.exp(-2j * np.pi * d[keep] / LAM)
    zero = np.zeros(1, np.float32)
    return dict(x=zero, y=zero, z=z, ant=np.array([[0.0, 0.0, H]], np.float32),
                rxo=np.zeros((1, 3), np.float32), spec=spec[None, :], S=S, d=d, bins=bins,
                keep=keep, want=want, lam=LAM, fs=FS, is_ti=is_ti)


# The reference rig's chirp (params.json, image_processor.py:359): 512 samples
# of 5.86116 MHz from 77.5001 GHz, bandwidth 3.0009 GHz, FFT to 4096 bins.
F0_RIG, SLOPE_RIG, NS_RIG, NFFT = 77.5001e9, 58.6116e6 * 1e6 / 10e6, 512, 4096
B_RIG = 3.0009e9


def chirp_rows(ant, target, rho=1.0 + 0.0j):
    """Monostatic chirp samples of a point scatterer at `target` seen from each
    row (FP64), in the sign convention the correlators undo:
    s_ij = rho exp(+j 2 pi (f0 + j slope) d_i / c), d_i = 2 |target - p_i|."""
    d = 2.0 * np.linalg.norm(ant.astype(np.float64) - np.asarray(target)[None, :], axis=1)
    f = F0_RIG + SLOPE_RIG * np.arange(NS_RIG)
    return rho * np.exp(2j * np.pi * d[:, None] * f[None, :] / C0), d


def geometric(x, n):
    """sum_{j<n} exp(j 2 pi j x) in closed form."""
    x = np.asarray(x, np.float64)
    num = 1.0 - np.exp(2j * np.pi * n * x)
    den = 1.0 - np.exp(2j * np.pi * x)
    small = np.abs(den) < 1e-12
    return np.where(small, n + 0j, num / np.where(small, 1.0, den))


def scatterer_case(n_ap=6, pitch=4e-3, height=0.3, rho=0.8 - 0.6j):
    """A point scatterer below an n_ap x n_ap monostatic aperture, its chirps
    zero-padded and FFT'd as image_processor.py:22-35 does (complex64 out).
    The FFT of the tone is S_i[k] = rho exp(j 2 pi f0 d_i / c) G(slope d_i / c -
    k / N) with G the geometric sum, so at the scatterer's voxel the image is
        I = rho sum_i exp(j 2 pi d_i (f0 / c - 1 / lambda)) G(slope d_i / c - b_i / N),
    b_i = floor(d_i / fs / 2), lambda the FP32 wavelength.  The geometry is
    chosen so that no d_i / fs / 2 is within 5 % of a bin edge (the FP32
    distance then cannot change a bin)."""
    lam = float(np.float32(C0 / F0_RIG))
    fs = float(np.float32(3e8 / (2 * B_RIG) * NS_RIG / NFFT))
    c = (np.arange(n_ap) - (n_ap - 1) / 2) * pitch
    for shift in np.arange(0, 0.02, 2.5e-4):
        ax, ay = np.meshgrid(c, c, indexing="ij")
        ant = np.stack([ax.ravel(), ay.ravel(), np.full(n_ap * n_ap, height + shift)], 1)
        ant = ant.astype(np.float32)
        target = np.array([np.float32(1.0e-3), np.float32(-2.0e-3), np.float32(0.0)], np.float64)
        s, d = chirp_rows(ant, target, rho)
        frac = d / fs / 2 - np.floor(d / fs / 2)
        if np.all((frac > 0.05) & (frac < 0.95)):
            break
    else:  # pragma: no cover
        raise AssertionError("no geometry clear of the bin edges")
    pad = np.zeros((len(ant), NFFT), np.complex64)
    pad[:, :NS_RIG] = s
    spec = np.fft.fft(pad.astype(np.complex128), axis=1).astype(np.complex64)
    b = np.floor(d / fs / 2).astype(np.int64)
    want = rho * np.sum(np.exp(2j * np.pi * d * (F0_RIG / C0 - 1.0 / lam)) *
                        geometric(SLOPE_RIG * d / C0 - b / NFFT, NS_RIG))
    return dict(ant=ant, target=target, spec=spec, d=d, bins=b, want=want, lam=lam, fs=fs,
                rho=rho)


def vote_single_row(phi=0.7, scale=2.5):
    """One bistatic row, one voxel, an image value I = scale c exp(j phi) not
    equal to the row's own contribution c = spec[b] exp(-j 2 pi d / lambda):
    the vote is (p + o / 2 - v) Re(c conj I) / |I| = (p + o / 2 - v) |c| cos(phi).
    The bin is read off d in FP64 (the geometry keeps d / fs / 2 away from
    an integer)."""
    p = np.array([0.01, -0.02, H], np.float32)
    o = np.array([0.0, -0.0069, 0.0], np.float32)
    v = np.array([0.003, 0.004, 0.1], np.float32)
    p64, o64, v64 = (a.astype(np.float64) for a in (p, o, v))
    d = np.linalg.norm(v64 - p64) + np.linalg.norm(v64 - p64 - o64)
    q = d / FS / 2
    assert 0.1 < q - np.floor(q) < 0.9
    b = int(np.floor(q))
    S = 128
    j = np.arange(S)
    spec = ((1.0 + 0.1 * j) * np.exp(1j * (0.5 + 0.21 * j))).astype(np.complex64)
    c = spec[b].astype(np.complex128) * np.exp(-2j * np.pi * d / LAM)
    img = np.complex64(scale * c * np.exp(1j * phi))
    vec = p64 + o64 / 2 - v64
    want = vec * (c * np.conj(np.complex128(img))).real / abs(np.complex128(img))
    return dict(x=v[:1], y=v[1:2], z=v[2:], ant=p[None, :], rxo=o[None, :], spec=spec[None, :],
                img=np.array([img], np.complex64), want=want, lam=LAM, fs=FS, c=c, vec=vec)


def vote_symmetric_4rx(n_tx=1):
    """One TX straight above the voxel and four RX offsets (+-2a, 0, 0),
    (0, +-2a, 0), a = 2^-8 m: the four round trips are equal in FP32 (the
    offsets only change signs), so the four rows (chirps of a point scatterer
    at the voxel, FFT'd) contribute the same c, and the vote's x and y terms
    cancel exactly: the unit normal is (0, 0, 1).  Returns the inputs."""
    a = 2.0 ** -8
    v = np.array([0.0, 0.0, 0.05], np.float32)
    p = np.array([[0.0, 0.0, 0.35]], np.float32)
    rxo = np.array([[2 * a, 0, 0], [-2 * a, 0, 0], [0, 2 * a, 0], [0, -2 * a, 0]], np.float32)
    rows = np.repeat(p, 4, axis=0)
    # bistatic chirp: d = |v - p| + |v - p - o|
    v64 = v.astype(np.float64)
    d = (np.linalg.norm(v64 - rows.astype(np.float64), axis=1) +
         np.linalg.norm(v64 - rows.astype(np.float64) - rxo.astype(np.float64), axis=1))
    assert np.all(d == d[0])
    f = F0_RIG + SLOPE_RIG * np.arange(NS_RIG)
    s = 0.9 * np.exp(2j * np.pi * d[:, None] * f[None, :] / C0)
    pad = np.zeros((4, NFFT), np.complex128)
    pad[:, :NS_RIG] = s
    spec = np.fft.fft(pad, axis=1).astype(np.complex64)
    lam = float(np.float32(C0 / F0_RIG))
    fs = float(np.float32(3e8 / (2 * B_RIG) * NS_RIG / NFFT))
    return dict(x=v[:1], y=v[1:2], z=v[2:], ant=rows, rxo=rxo, spec=spec, lam=lam, fs=fs)
