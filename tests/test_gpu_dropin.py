"""The NumPy drop-ins (image_cuda, normal_estimates_cuda) against the reference's
host semantics restated over the oracle: layout, double normalisation,
antenna groups, NaN scatter and phase centres."""
import numpy as np
import pytest

from mmnorm_amd import py_image_gpu, py_sn_est_gpu, synth

pytestmark = pytest.mark.gpu
C0 = synth.C0


def real_like_inputs(seed=7, n_meas=20, S=64, spectra=False):
    """Real-data layout: (n_meas, samples, 4 RX), 'almost-square' RX offsets, TI radar."""
    rng = np.random.default_rng(seed)
    lam = C0 / 77.5001e9
    rxo = np.array([[-0.005 - lam / 2 * k, 0, 0] for k in (3, 2, 1, 0)])
    locs = np.stack([rng.uniform(-0.03, 0.03, n_meas), rng.uniform(-0.03, 0.03, n_meas),
                     rng.uniform(0.3, 0.32, n_meas)], 1)
    if spectra:
        S = 4096
    data = (rng.normal(size=(n_meas, S, 4)) + 1j * rng.normal(size=(n_meas, S, 4))).astype(
        np.complex64)
    xl, yl, zl = np.arange(-0.01, 0.01, 0.002), np.arange(-0.008, 0.008, 0.002), \
        np.arange(-0.006, 0.004, 0.002)
    return xl, yl, zl, locs, data, rxo, lam


def _layout(locs, data, R):
    meas = np.transpose(data, (0, 2, 1)).reshape(-1, data.shape[1])
    return np.repeat(locs, R, axis=0), meas


def test_image_cuda_full_ref_variant(oracle):
    xl, yl, zl, locs, data, rxo, lam = real_like_inputs()
    slope = 58.6116 * 1e6 * 1e6 / 10e6
    img, centers = py_image_gpu.image_cuda(xl, yl, zl, locs, data, rxo, slope, lam, 3.0009e9,
                                           512, use_4_rx=True, is_ti_radar=True,
                                           use_interpolated_processing=False, variant="ref")
    ant, meas = _layout(locs, data, 4)
    raw = oracle.bp_ref(xl, yl, zl, ant, meas, rxo, slope, lam, is_ti=True)
    want = raw.reshape(len(xl), len(yl), len(zl)).astype(np.complex64)
    want /= len(locs)
    want /= len(locs)
    assert img.dtype == np.complex64 and img.shape == want.shape
    assert np.max(np.abs(img - want)) <= 1e-6 * np.max(np.abs(want))
    wc = (ant + (ant + np.tile(rxo, (len(locs), 1)))) / 2
    assert centers.dtype == np.float64
    np.testing.assert_array_equal(centers, wc)


@pytest.mark.parametrize("variant", ["rec", "profile"])
def test_image_cuda_full_variant_close_to_fp64(oracle, variant):
    xl, yl, zl, locs, data, rxo, lam = real_like_inputs(seed=8)
    slope = 5.86116e6
    img = py_image_gpu.image_cuda(xl, yl, zl, locs, data, rxo, slope, lam, 3.0009e9, 512,
                                  use_4_rx=True, is_ti_radar=True,
                                  return_phase_centers=False, variant=variant)
    ant, meas = _layout(locs, data, 4)
    f0 = C0 / float(np.float32(lam))
    want = oracle.bp_f64(xl, yl, zl, ant, meas, rxo, f0, float(np.float32(slope)),
                         is_ti=True).reshape(img.shape) / len(locs) ** 2
    assert np.linalg.norm(img - want) / np.linalg.norm(want) < 1e-5


@pytest.mark.parametrize("variant", ["rec", "ref"])
def test_image_cuda_interpolated_groups(oracle, monkeypatch, variant):
    """Antenna groups normalised by their own TX count, then summed
    (py_image_gpu.py:339-362); the default (MI355X) and the reference-arithmetic
    kernels against the restatement of imaging_gpu.cu."""
    monkeypatch.setattr(py_image_gpu, "MAX_GROUP_ROWS", 32)  # 8 TX per group with 4 RX
    xl, yl, zl, locs, data, rxo, lam = real_like_inputs(seed=9, n_meas=19, spectra=True)
    fs = np.float32(3e8 / (2 * 3.0009e9) * 512 / 4096)
    img, centers = py_image_gpu.image_cuda(xl, yl, zl, locs, data, rxo, 5.86116e6, lam,
                                           3.0009e9, 512, use_4_rx=True, is_ti_radar=True,
                                           use_interpolated_processing=True, variant=variant)
    want = None
    for lo in range(0, 19, 8):
        ant, meas = _layout(locs[lo:lo + 8], data[lo:lo + 8], 4)
        raw = oracle.bp_ref(xl, yl, zl, ant, meas, rxo, 5.86116e6, lam, fs, is_ti=True,
                            interp=True).reshape(len(xl), len(yl), len(zl))
        part = raw.astype(np.complex64)
        part /= len(locs[lo:lo + 8])
        want = part if want is None else want + part
    assert np.max(np.abs(img - want)) <= 1e-6 * np.max(np.abs(want))
    assert centers.shape == (19 * 4, 3)


@pytest.mark.parametrize("variant", ["fast", "ref"])
def test_normal_estimates_interpolated(oracle, variant):
    xl, yl, zl, locs, data, rxo, lam = real_like_inputs(seed=10, spectra=True)
    img = py_image_gpu.image_cuda(xl, yl, zl, locs, data, rxo, 5.86116e6, lam, 3.0009e9, 512,
                                  use_4_rx=True, is_ti_radar=True,
                                  use_interpolated_processing=True, return_phase_centers=False)
    nrm = py_sn_est_gpu.normal_estimates_cuda(img, xl, yl, zl, locs, data, rxo, 5.86116e6,
                                              lam, 3.0009e9, 512, use_4_rx=True,
                                              is_ti_radar=True, initial_filter_percent=0.3,
                                              variant=variant)
    filt = np.abs(img) > np.max(np.abs(img)) * 0.3
    assert nrm.dtype == np.float64 and nrm.shape == img.shape + (3,)
    assert np.all(np.isnan(nrm[~filt])) and not np.any(np.isnan(nrm[filt]))
    ant, meas = _layout(locs, data, 4)
    fs = np.float32(3e8 / (2 * 3.0009e9) * 512 / 4096)
    want = oracle.sn_ref(np.flatnonzero(filt), img.ravel(), xl, yl, zl, ant, meas, rxo, lam,
                         fs, is_ti=True)
    assert np.max(np.abs(nrm[filt] - want)) <= 1e-5 * np.max(np.abs(want))


@pytest.mark.parametrize("variant", ["fast", "profile"])
def test_normal_estimates_full_frequency(oracle, variant):
    xl, yl, zl, locs, data, rxo, lam = real_like_inputs(seed=12)
    img = py_image_gpu.image_cuda(xl, yl, zl, locs, data, rxo, 5.86116e6, lam, 3.0009e9, 512,
                                  use_4_rx=True, is_ti_radar=True, return_phase_centers=False,
                                  variant="rec" if variant == "fast" else "profile")
    nrm = py_sn_est_gpu.normal_estimates_cuda(img, xl, yl, zl, locs, data, rxo, 5.86116e6,
                                              lam, 3.0009e9, 512, use_4_rx=True,
                                              is_ti_radar=True,
                                              use_interpolated_processing=False,
                                              variant=variant)
    filt = np.abs(img) > np.max(np.abs(img)) * 0.05
    ant, meas = _layout(locs, data, 4)
    f0 = C0 / float(np.float32(lam))
    wi, wm = oracle.bp_f64(xl, yl, zl, ant, meas, rxo, f0, float(np.float32(5.86116e6)),
                           is_ti=True, vidx=np.flatnonzero(filt), moment=True)
    I = img.ravel()[np.flatnonzero(filt)].astype(np.complex128)
    wn = (I.real[:, None] * wm.real + I.imag[:, None] * wm.imag) / np.abs(I)[:, None]
    got = nrm[filt]
    cos = np.sum(got * wn, 1) / np.linalg.norm(got, axis=1) / np.linalg.norm(wn, axis=1)
    assert np.degrees(np.arccos(np.clip(cos.min(), -1, 1))) < 0.05
    assert np.all(np.isnan(nrm[~filt]))


@pytest.mark.parametrize("shape3", [False, True])
@pytest.mark.parametrize("variant", ["fast", "ref"])
def test_single_rx_layout_use_4_rx_false(oracle, variant, shape3):
    """use_4_rx=False (py_sn_est_gpu.py:276-278, py_image_gpu.py:228-235): the
    measurements are flattened and read as len(antenna_locs) rows of
    radar_data.shape[1] samples, the offsets wrapped in a list; both a
    (meas, samples) and a (meas, samples, 1) array, one RX offset, the
    interpolated image and vote against the oracle."""
    xl, yl, zl, locs, data, _, lam = real_like_inputs(seed=13, n_meas=24, spectra=True)
    d1 = data[:, :, 1:2] if shape3 else data[:, :, 1]
    rxo = [[-0.0051, 0.0, 0.0]]
    fs = np.float32(3e8 / (2 * 3.0009e9) * 512 / 4096)
    img = py_image_gpu.image_cuda(xl, yl, zl, locs, d1, rxo, 5.86116e6, lam, 3.0009e9, 512,
                                  use_4_rx=False, is_ti_radar=True,
                                  use_interpolated_processing=True, return_phase_centers=False)
    meas = data[:, :, 1]
    raw = oracle.bp_ref(xl, yl, zl, locs, meas, np.array(rxo), 5.86116e6, lam, fs, is_ti=True,
                        interp=True).reshape(img.shape).astype(np.complex64)
    raw /= len(locs)
    assert np.max(np.abs(img - raw)) <= 1e-6 * np.max(np.abs(raw))
    nrm = py_sn_est_gpu.normal_estimates_cuda(img, xl, yl, zl, locs, d1, rxo, 5.86116e6, lam,
                                              3.0009e9, 512, use_4_rx=False, is_ti_radar=True,
                                              initial_filter_percent=0.3, variant=variant)
    filt = np.abs(img) > np.max(np.abs(img)) * 0.3
    want = oracle.sn_ref(np.flatnonzero(filt), img.ravel(), xl, yl, zl, locs, meas,
                         np.array(rxo), lam, fs, is_ti=True)
    assert np.all(np.isnan(nrm[~filt]))
    assert np.max(np.abs(nrm[filt] - want)) <= 1e-5 * np.max(np.abs(want))
