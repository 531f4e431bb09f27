"""The HIP kernels of the interpolated branch and its vote against the closed
forms of tests/closed_forms.py (the same cases that pin the CPU oracle):
exact bins and the skip rule on hand-constructed distances, the FFT'd point
scatterer's focus value, a single row's vote and a symmetric four-RX
aperture's normal.  Kernels: k_image_mirror (the reference's arithmetic),
k_bp_interp (MI355X form, vote moment fused), k_sn_mirror and
k_sn_from_moment."""
import numpy as np
import pytest
import torch

import closed_forms as cf
from mmnorm_amd import ops

pytestmark = pytest.mark.gpu


def T(a, dt):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(torch.device("cuda", 0))


def _args(c, spec=None):
    return (T(c["spec"] if spec is None else spec, np.complex64), T(c["ant"], np.float32),
            T(c["x"], np.float32), T(c["y"], np.float32), T(c["z"], np.float32))


@pytest.mark.parametrize("is_ti", [False, True])
@pytest.mark.parametrize("kernel", ["mirror", "fast"])
def test_interp_bins_and_skip(kernel, is_ti):
    c = cf.bin_case(is_ti=is_ti)
    a = _args(c)
    if kernel == "mirror":
        got = ops.cuda_image(*a, T(c["rxo"], np.float32), slope=0.0, wavelength=c["lam"],
                             fft_spacing=c["fs"], is_ti=is_ti, interpolated=True)
    else:
        got = ops.backproject_interp(*a, T(c["rxo"], np.float32), wavelength=c["lam"],
                                     fft_spacing=c["fs"], is_ti=is_ti)
    got = got.cpu().numpy().ravel()
    assert np.all(got[~c["keep"]] == 0)
    mag = np.abs(c["spec"][0]).astype(np.float64)
    bins = np.array([np.argmin(np.abs(mag - abs(g))) for g in got[c["keep"]]])
    np.testing.assert_array_equal(bins, c["bins"][c["keep"]])
    tol = 2e-4 if is_ti else 2e-6
    err = np.abs(got[c["keep"]] - c["want"][c["keep"]]) / np.abs(c["want"][c["keep"]])
    assert err.max() < tol


@pytest.mark.parametrize("kernel", ["mirror", "fast"])
def test_interp_point_scatterer_focus(kernel):
    c = cf.scatterer_case()
    t = c["target"].astype(np.float32)
    a = (T(c["spec"], np.complex64), T(c["ant"], np.float32), T(t[:1], np.float32),
         T(t[1:2], np.float32), T(t[2:], np.float32))
    if kernel == "mirror":
        got = ops.cuda_image(*a, None, slope=0.0, wavelength=c["lam"], fft_spacing=c["fs"],
                             interpolated=True)
    else:
        got = ops.backproject_interp(*a, wavelength=c["lam"], fft_spacing=c["fs"])
    got = complex(got.cpu().numpy().ravel()[0])
    assert abs(got - c["want"]) < 2e-4 * abs(c["want"])


@pytest.mark.parametrize("phi", [0.7, 2.5])
@pytest.mark.parametrize("kernel", ["mirror", "moment"])
def test_vote_single_row(kernel, phi):
    c = cf.vote_single_row(phi=phi)
    a = _args(c)
    img = T(c["img"], np.complex64)
    vidx = T(np.array([0]), np.int32)
    if kernel == "mirror":
        n = ops.sn_est(img, vidx, *a, T(c["rxo"], np.float32), wavelength=c["lam"],
                       fft_spacing=c["fs"])
    else:
        _, mom = ops.backproject_interp(*a, T(c["rxo"], np.float32), wavelength=c["lam"],
                                        fft_spacing=c["fs"], vox_idx=vidx, moment=True)
        n = ops.sn_from_moment(img, mom, vidx)
    got = n.cpu().numpy()[0].astype(np.float64)
    np.testing.assert_allclose(got, c["want"], rtol=2e-4, atol=0)


@pytest.mark.parametrize("kernel", ["mirror", "moment"])
def test_vote_symmetric_four_rx(kernel):
    c = cf.vote_symmetric_4rx()
    a = _args(c)
    rx = T(c["rxo"], np.float32)
    img = ops.backproject_interp(*a, rx, wavelength=c["lam"], fft_spacing=c["fs"]).reshape(-1)
    vidx = T(np.array([0]), np.int32)
    if kernel == "mirror":
        n = ops.sn_est(img.contiguous(), vidx, *a, rx, wavelength=c["lam"], fft_spacing=c["fs"])
    else:
        _, mom = ops.backproject_interp(*a, rx, wavelength=c["lam"], fft_spacing=c["fs"],
                                        vox_idx=vidx, moment=True)
        n = ops.sn_from_moment(img.contiguous(), mom, vidx)
    got = n.cpu().numpy()[0]
    assert got[0] == 0.0 and got[1] == 0.0 and got[2] > 0
