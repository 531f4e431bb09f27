"""Closed-form pins of the oracle's interpolated branch (or_bp_ref with
interp, imaging_gpu.cu:90-105) and of its vote (or_sn_ref,
surface_normal_est.cu:110-149), independent of the oracle's expressions:
hand-constructed distances whose bins are integer arithmetic, the skip rule
at d = fs * S, an FFT'd point scatterer whose focus value is a geometric
sum, a single row's vote direction and weight, and a symmetric aperture's
normal (0, 0, 1).  A wrong bin formula (d / fs, round instead of floor, a
missing / 2), a wrong skip test (>= for >), a wrong phase sign or
wavelength, a flipped vote vector or o instead of o / 2, or Im for Re in the
weight each fail one of them."""
import numpy as np
import pytest

import closed_forms as cf


@pytest.mark.parametrize("is_ti", [False, True])
def test_interp_bins_and_skip_on_exact_distances(oracle, is_ti):
    c = cf.bin_case(is_ti=is_ti)
    got = oracle.bp_ref(c["x"], c["y"], c["z"], c["ant"], c["spec"], c["rxo"], 0.0, c["lam"],
                        c["fs"], is_ti=is_ti, interp=True)
    # skipped rows contribute exactly nothing
    assert np.all(got[~c["keep"]] == 0)
    # every kept voxel reads the bin the integer arithmetic gives: the
    # spectrum's magnitudes are distinct, so |I| identifies the bin
    mag = np.abs(c["spec"][0]).astype(np.float64)
    got_bins = np.array([np.argmin(np.abs(mag - abs(g))) for g in got[c["keep"]]])
    np.testing.assert_array_equal(got_bins, c["bins"][c["keep"]])
    # and the phase is exp(-j 2 pi d / lambda) of that distance
    tol = 2e-4 if is_ti else 2e-6    # FP32 round trip of d + 0.15: ~3e-8 m of phase
    err = np.abs(got[c["keep"]] - c["want"][c["keep"]]) / np.abs(c["want"][c["keep"]])
    assert err.max() < tol
    if not is_ti:
        assert c["keep"].sum() >= 10 and (~c["keep"]).sum() == 2
        # the edge: d == fs * S is kept and reads bin S / 2
        assert c["bins"][c["keep"]].max() == c["S"] // 2


def test_interp_point_scatterer_fft_closed_form(oracle):
    c = cf.scatterer_case()
    t = c["target"].astype(np.float32)
    got = oracle.bp_ref(t[:1], t[1:2], t[2:], c["ant"], c["spec"], np.zeros((1, 3)), 0.0,
                        c["lam"], c["fs"], interp=True)[0]
    assert abs(got - c["want"]) < 2e-4 * abs(c["want"])
    # it focuses: the rows add nearly in phase, each near the Dirichlet peak
    n_rows = len(c["ant"])
    assert abs(got) > 0.85 * abs(c["rho"]) * n_rows * cf.NS_RIG
    # and the focus is the brightest voxel across the aperture's footprint
    off = np.arange(-4, 5) * 6e-3
    xs = (t[0] + off).astype(np.float32)
    ys = (t[1] + off).astype(np.float32)
    img = oracle.bp_ref(xs, ys, t[2:], c["ant"], c["spec"], np.zeros((1, 3)), 0.0, c["lam"],
                        c["fs"], interp=True).reshape(9, 9)
    assert np.unravel_index(np.argmax(np.abs(img)), img.shape) == (4, 4)


@pytest.mark.parametrize("phi", [0.7, 2.5])
def test_vote_single_row_direction_and_weight(oracle, phi):
    c = cf.vote_single_row(phi=phi)
    got = oracle.sn_ref([0], c["img"], c["x"], c["y"], c["z"], c["ant"], c["spec"], c["rxo"],
                        c["lam"], c["fs"])[0].astype(np.float64)
    np.testing.assert_allclose(got, c["want"], rtol=2e-4, atol=0)
    # the unit vote points from the voxel to the virtual antenna p + o/2 (or
    # away from it when the row's phase opposes the image's)
    u = c["vec"] / np.linalg.norm(c["vec"])
    assert np.dot(got / np.linalg.norm(got), u) == pytest.approx(np.sign(np.cos(phi)), abs=1e-6)
    assert np.linalg.norm(got) == pytest.approx(
        np.linalg.norm(c["vec"]) * abs(c["c"]) * abs(np.cos(phi)), rel=2e-4)


def test_vote_symmetric_four_rx_is_vertical(oracle):
    c = cf.vote_symmetric_4rx()
    img = oracle.bp_ref(c["x"], c["y"], c["z"], c["ant"], c["spec"], c["rxo"], 0.0, c["lam"],
                        c["fs"], interp=True)
    n = oracle.sn_ref([0], img, c["x"], c["y"], c["z"], c["ant"], c["spec"], c["rxo"],
                      c["lam"], c["fs"])[0]
    assert n[0] == 0.0 and n[1] == 0.0 and n[2] > 0


def test_closed_forms_catch_planted_slips():
    """The closed forms above are sharp: restatements with a planted slip
    (here in NumPy, on the CPU) disagree with them."""
    c = cf.bin_case()
    d, S = c["d"], c["S"]
    good = np.floor(d.astype(np.float32) / np.float32(cf.FS) / np.float32(2)).astype(int)
    assert np.array_equal(good, c["bins"])
    assert not np.array_equal(np.floor(d / cf.FS).astype(int), c["bins"])        # no / 2
    assert not np.array_equal(np.rint(d / cf.FS / 2).astype(int), c["bins"])     # round
    assert not np.array_equal(d < cf.FS * S, c["keep"])                          # >= skip
    v = cf.vote_single_row()
    flipped = -v["vec"] * (v["c"] * np.conj(np.complex128(v["img"][0]))).real
    assert not np.allclose(flipped / abs(v["img"][0]), v["want"])
    im = v["vec"] * (v["c"] * np.conj(np.complex128(v["img"][0]))).imag / abs(v["img"][0])
    assert not np.allclose(im, v["want"], rtol=1e-2)
