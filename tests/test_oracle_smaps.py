"""Pins of the coil-map oracle (CPU): window values at hand-picked points,
the blur of a constant and of a delta, LSQR against lstsq and its first
iterate's closed form, unit root sum of squares."""

import numpy as np
import pytest

from helpers import rand_cplx
from oracle import smaps


def test_window_closed_forms():
    k = np.array([[0.05, 0.1], [0.09, 0.15], [0.1, 0.0], [0.0, 0.0]])
    np.testing.assert_array_equal(smaps.window(k, (0.1, 0.2), "ellipse"), [1, 0, 1, 1])
    k = np.array([[0.1, -0.1], [0.11, 0.0], [0.0, 0.05]])
    np.testing.assert_array_equal(smaps.window(k, 0.1, "rect"), [1, 0, 1])
    k = np.array([[0.0, 0.0], [0.1, 0.0], [0.03, 0.04]])
    np.testing.assert_allclose(smaps.window(k, 0.1, "hann"), [1.0, 0.0, 0.5], atol=1e-15)
    np.testing.assert_allclose(smaps.window(k[:2], 0.1, "hamming"), [1.0, 2 * 0.53836 - 1])


def test_blur_constant_and_delta():
    c = np.full((2, 9, 11), 3.0 - 1.0j)
    np.testing.assert_allclose(smaps.blur(c, (1.3, 0.8)), c, rtol=1e-12)
    sigma = 1.5
    r = int(4.0 * sigma + 0.5)
    j = np.arange(-r, r + 1)
    w = np.exp(-0.5 * j ** 2 / sigma ** 2)
    w /= w.sum()
    d = np.zeros((1, 31))
    d[0, 15] = 1.0
    out = smaps.blur(d, (sigma,)).real[0]
    np.testing.assert_allclose(out[15 - r:15 + r + 1], w, rtol=1e-12)
    assert np.all(out[:15 - r] == 0) and np.all(out[15 + r + 1:] == 0)


def test_lsqr_converges_and_first_iterate():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((40, 10)) + 1j * rng.standard_normal((40, 10))
    b = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    x = smaps.lsqr(A, b, 100)
    np.testing.assert_allclose(x, np.linalg.lstsq(A, b, rcond=None)[0], rtol=1e-6, atol=1e-8)
    g = A.conj().T @ b
    x1 = smaps.lsqr(A, b, 1)
    np.testing.assert_allclose(x1, (np.vdot(g, g) / np.vdot(A @ g, A @ g)) * g, rtol=1e-10)


def test_rss_normalize_unit():
    m = rand_cplx((4, 6, 7), 1)
    n = smaps.rss_normalize(m)
    np.testing.assert_allclose(np.linalg.norm(n, axis=0), 1.0, rtol=1e-12)
