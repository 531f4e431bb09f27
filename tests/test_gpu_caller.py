"""The GPU-resident caller (image_processor.generate_mmwave_surface_normals)
against the same steps composed from the NumPy drop-ins."""
import numpy as np
import pytest
import torch

from mmnorm_amd import image_processor, py_image_gpu, py_sn_est_gpu

pytestmark = pytest.mark.gpu

GRID = dict(min_x=0.05, max_x=0.066, min_y=0.17, max_y=0.182, min_z=-0.44, max_z=-0.43,
            step=0.001)


def inputs(seed, n=24, samples=64):
    rng = np.random.default_rng(seed)
    poses = np.stack([0.42 + 0.002 * (np.arange(n) % 6), -0.23 + 0.15 + 0.002 * (np.arange(n) // 6),
                      0.23 * np.ones(n), np.zeros(n)], 1)
    data = (rng.normal(size=(n, samples + 8, 4)) + 1j * rng.normal(size=(n, samples + 8, 4)))
    return poses, data.astype(np.complex64)


@pytest.mark.parametrize("interp", [True, False])
def test_caller_matches_composed_dropins(interp):
    poses, data = inputs(3)
    min_f, ns, bw = 77e9, 64, 3.0009e9
    img, axes, nrm, centers, meas = image_processor.generate_mmwave_surface_normals(
        poses, data, "almost-square", min_f, ns, bw, use_interpolated_processing=interp,
        grid=GRID)
    # the same steps through the NumPy drop-ins
    locs = poses[:, :3] - np.array(image_processor.LOCATION0)
    lam = 2.99792458e8 / min_f
    rxo = image_processor.rx_offsets_for("almost-square", lam)
    clip = data[:, :ns, :4]
    md = np.fft.fft(np.concatenate([clip, np.zeros((len(clip), 4096 - ns, 4), np.complex64)], 1),
                    axis=1) if interp else clip
    xl, yl, zl = image_processor.imaging_grid(GRID)
    ref_img, ref_c = py_image_gpu.image_cuda(xl, yl, zl, locs, md, rxo,
                                             image_processor.SLOPE_REAL, lam, bw, ns,
                                             use_4_rx=True, is_ti_radar=True,
                                             use_interpolated_processing=interp)
    ref_n = py_sn_est_gpu.normal_estimates_cuda(ref_img, xl, yl, zl, locs, md, rxo,
                                                image_processor.SLOPE_REAL, lam, bw, ns,
                                                use_4_rx=True, is_ti_radar=True,
                                                use_interpolated_processing=interp)
    ref_n = ref_n / np.linalg.norm(ref_n, axis=-1, keepdims=True)
    assert np.max(np.abs(img - ref_img)) <= 1e-4 * np.max(np.abs(ref_img))
    assert np.array_equal(np.isnan(nrm), np.isnan(ref_n))
    ok = ~np.isnan(nrm[..., 0])
    cosang = np.sum(nrm[ok] * ref_n[ok], -1)
    assert np.degrees(np.arccos(np.clip(cosang.min(), -1, 1))) < 0.01
    assert np.allclose(centers, ref_c) and np.array_equal(meas, clip)
    assert all(np.array_equal(a, b) for a, b in zip(axes, (xl, yl, zl)))


def test_interpolate_measurements_matches_numpy():
    _, data = inputs(4, n=5, samples=100)
    got = image_processor.interpolate_measurements(data[:, :100, :])
    pad = np.zeros((5, 4096, 4), np.complex64)
    pad[:, :100] = data[:, :100]
    want = np.fft.fft(pad, axis=1)
    assert np.max(np.abs(got - want)) <= 1e-4 * np.max(np.abs(want))
