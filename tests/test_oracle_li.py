"""The restatement of scikit-image 0.24.0's threshold_li (oracle.py_skimage_li)
pinned by closed forms -- constant and two-level images, three levels where
the half-gap tolerance stops the iteration after one update or after two --
NaN handling, the callback limit of sdf.py:53-58, and agreement with the
FP64 fixed-point iteration (or_li_threshold) to float32 precision."""
import numpy as np
import pytest

from oracle import oracle as O


def f32(*v):
    return np.array(v, np.float32)


def test_constant_and_two_levels():
    thr, calls = O.py_skimage_li(np.full(100, 7.5, np.float32))
    assert thr == np.float32(7.5) and calls == 0
    # two levels: the background mean is 0 at once, the threshold is the mean
    img = np.repeat(f32(5.0, 9.0), [300, 100])
    thr, calls = O.py_skimage_li(img)
    assert calls == 1
    assert thr == np.float32(5.0) + np.mean(img - np.float32(5.0))
    assert abs(float(thr) - 6.0) < 1e-6


def test_three_levels_tolerance_stops_after_one_update():
    # x = {0, 1, 10}: t0 = 11/3; one update gives 9.5 / ln 20, a step of 0.4956,
    # below the tolerance 0.5 (half the gap 1): two callbacks, where the exact
    # fixed-point iteration needs a second, confirming update
    img = np.tile(f32(2.0, 3.0, 12.0), 1000)
    thr, calls = O.py_skimage_li(img)
    assert calls == 2
    assert abs(float(thr) - (2.0 + 9.5 / np.log(20.0))) < 2e-6
    t64, k64 = O.py_li_threshold(img, 1.0)
    assert k64 == 2 and abs(float(thr) - t64) < 2e-6


def test_three_levels_two_updates():
    # x = {0, 0.5, 10}: tolerance 0.25; 10.5/3 -> 9.75 / ln 40 (a step of 0.86),
    # then the same partition gives the same value: three callbacks
    img = np.tile(f32(2.0, 2.5, 12.0), 1000)
    thr, calls = O.py_skimage_li(img)
    assert calls == 3
    assert abs(float(thr) - (2.0 + 9.75 / np.log(40.0))) < 2e-6


def test_nans_dropped_and_callback_limit():
    rng = np.random.default_rng(4)
    img = (np.abs(rng.normal(size=5000)) ** 2).astype(np.float32)
    with_nan = img.copy()
    with_nan[::7] = np.nan
    thr, calls = O.py_skimage_li(with_nan)
    want, want_calls = O.py_skimage_li(img[np.arange(img.size) % 7 != 0])
    assert thr == want and calls == want_calls and calls > 2
    full, full_calls = O.py_skimage_li(img)
    with pytest.raises(O.LiAbort):
        O.py_skimage_li(img, max_callbacks=full_calls - 1)
    assert O.py_skimage_li(img, max_callbacks=full_calls) == (full, full_calls)


@pytest.mark.parametrize("ex", [2, 2.5, 3.5, 4])
def test_agrees_with_fp64_iteration(oracle, ex):
    rng = np.random.default_rng(int(ex * 10))
    n = 100_000
    z = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    z[: n // 50] *= 6
    mask, thr, calls, abs_img = O.skimage_li_mask(z, ex)
    t64, k64 = oracle.li_threshold(z, ex)
    assert abs(float(thr) - t64) <= 1e-5 * t64
    assert calls - 1 <= k64
    a64 = np.abs(z.astype(np.complex128)) ** ex
    diff = mask != (a64 < t64)
    assert np.all(np.abs(a64[diff] - t64) <= 1e-5 * t64)
