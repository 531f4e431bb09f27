"""The image drop-in's normalisation (py_image_gpu.py:301-302 and :330): the
product's fast path (one float32 multiply per component, done where the image
lies) and its fallback against NumPy's own complex64 steps, bit for bit, on
CPU tensors."""
import numpy as np
import pytest
import torch

from mmnorm_amd import py_image_gpu


def numpy_steps(raw, n, divisions):
    """The reference's lines as written: p = re + 1j*im, then p /= n."""
    p = raw.real + 1j * raw.imag
    p = p.astype(np.complex64, copy=False)
    for _ in range(divisions):
        p /= n
    return p


def images(rng, size):
    """Typical magnitudes, a wide dynamic range, and one with the special values."""
    typ = rng.standard_normal(2 * size).astype(np.float32) * np.float32(1e3)
    wide = (rng.standard_normal(2 * size) * 10.0 ** rng.integers(-30, 30, 2 * size)).astype(
        np.float32)
    edge = typ.copy()
    k = rng.permutation(2 * size)[:70]
    edge[k[:10]] = 0.0
    edge[k[10:20]] = -0.0
    edge[k[20:30]] = np.inf
    edge[k[30:40]] = -np.inf
    edge[k[40:50]] = np.nan
    edge[k[50:60]] = np.float32(1e-45)
    edge[k[60:]] = np.float32(-2e-38)
    return {"typical": typ, "wide": wide, "edge": edge}


@pytest.mark.parametrize("n", [1, 7, 3600, 10000, 2 ** 25 + 3])
@pytest.mark.parametrize("divisions", [1, 2])
def test_scaling_equals_numpy_bit_for_bit(n, divisions):
    rng = np.random.default_rng(n + divisions)
    for name, flat in images(rng, 4096).items():
        raw = flat.view(np.complex64).reshape(16, 16, 16)
        want = numpy_steps(raw, n, divisions)
        got = py_image_gpu._reference_scaled(torch.from_numpy(raw.copy()), n, divisions)
        assert got.dtype == np.complex64 and got.shape == want.shape, name
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, n, divisions)


def test_fast_path_taken_only_where_it_is_exact():
    """Typical images go through the multiply (the tensor is scaled in place);
    ones with a zero, a NaN, or values that would turn subnormal keep NumPy's
    steps (the tensor is left as it was)."""
    rng = np.random.default_rng(0)
    typ = images(rng, 512)["typical"].view(np.complex64)
    t = torch.from_numpy(typ.copy())
    py_image_gpu._reference_scaled(t, 3600, 2)
    assert not np.array_equal(t.numpy(), typ)
    for bad in (0.0, np.nan, np.inf, 1e-34):
        a = typ.copy()
        a.real[5] = bad
        t = torch.from_numpy(a.copy())
        got = py_image_gpu._reference_scaled(t, 3600, 2)
        assert np.array_equal(t.numpy().view(np.uint32), a.view(np.uint32)), bad
        assert np.array_equal(got.view(np.uint32), numpy_steps(a, 3600, 2).view(np.uint32))
