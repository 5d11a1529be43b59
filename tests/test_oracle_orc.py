"""Pins of the off-resonance oracle (oracle/orc.py) and of the host-side
field-map factorizations (hipnufft.field_map, torch on the CPU here)."""

import numpy as np
import pytest

from helpers import gauss_smaps, rand_cplx, rel_l2, truncnorm_traj
from oracle import ndft
from oracle import orc


def piecewise_field(shape, levels, seed):
    """A field map taking only the values ``levels`` and its level index."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(levels), shape)
    return np.asarray(levels)[idx], idx


def exact_interpolators(levels, idx, t):
    """``exp(t f) = sum_l exp(t w_l) [f == w_l]``: an exact factorization."""
    B = np.exp(np.outer(t, levels))
    C = np.stack([(idx == l).astype(np.complex128) for l in range(len(levels))])
    return B, C


LEVELS = np.array([2j * np.pi * -120, 2j * np.pi * 35, 30 + 2j * np.pi * 180, 2j * np.pi * 260])


@pytest.mark.parametrize("shape,n_shots", [((10, 12), 1), ((10, 12), 3), ((6, 5, 4), 2)])
@pytest.mark.parametrize("sense", [True, False])
def test_exact_factorization_equals_extended_matrix(shape, n_shots, sense):
    NK = 40
    d = len(shape)
    traj = truncnorm_traj(NK * n_shots, d, 1)
    t = np.linspace(0, 8e-3, NK)
    f, idx = piecewise_field(shape, LEVELS, 2)
    B, C = exact_interpolators(LEVELS, idx, t)
    E = orc.extended_fourier_matrix(traj, shape, f, np.tile(t, n_shots))
    nb, nc = 2, 3
    smaps = gauss_smaps(nc, shape, 3) if sense else None
    x = rand_cplx((nb, 1 if sense else nc, *shape), 4).astype(np.complex128)
    y = orc.op(traj, shape, x, B, C, smaps)
    imgs = (smaps[None] * x if sense else x).reshape(nb, nc, -1)
    ref = imgs @ E.T
    assert rel_l2(y, ref) < 1e-12
    k = rand_cplx((nb, nc, NK * n_shots), 5).astype(np.complex128)
    dens = np.random.default_rng(6).uniform(0.5, 1.5, NK * n_shots)
    got = orc.adj(traj, shape, k, B, C, smaps, dens)
    back = (k * dens) @ E.conj()
    if sense:
        back = np.sum(np.conj(smaps.reshape(nc, -1))[None] * back, axis=1, keepdims=True)
    assert rel_l2(got.reshape(back.shape), back) < 1e-12


def test_adjointness():
    shape, NK, L = (8, 9), 30, 3
    traj = truncnorm_traj(NK * 2, 2, 7)
    B = rand_cplx((NK, L), 8).astype(np.complex128)
    C = rand_cplx((L, *shape), 9).astype(np.complex128)
    smaps = gauss_smaps(2, shape, 10)
    x = rand_cplx((1, 1, *shape), 11)
    y = rand_cplx((1, 2, NK * 2), 12)
    lhs = np.vdot(orc.op(traj, shape, x, B, C, smaps), y)
    rhs = np.vdot(x, orc.adj(traj, shape, y, B, C, smaps))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_one_interpolator_is_the_plain_ndft():
    shape = (8, 8)
    traj = truncnorm_traj(50, 2, 13)
    x = rand_cplx((1, 1, *shape), 14)
    y = orc.op(traj, shape, x, np.ones((50, 1)), np.ones((1, *shape)))
    assert rel_l2(y.ravel(), ndft.type2(traj, x[0, 0], shape) / ndft.norm_factor(shape)) < 1e-12


def _maps(N, complex_):
    from hipnufft import make_b0map, make_t2smap

    b0, mask = make_b0map((N, N))
    r2s = None
    if complex_:
        t2s, _ = make_t2smap((N, N), 15.0)
        r2s = mask * (1.0 / (t2s * 1e-3 + 1e-9))
    return orc.complex_fieldmap(b0, r2s), mask


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("method", ["svd", "mti", "mfi"])
@pytest.mark.parametrize("L", [40, -1])
def test_oracle_factorizations_reproduce_the_exponential(method, L, complex_):
    """The reference's own bound (test_offres_exp_approx.py:98-101)."""
    f, mask = _maps(48, complex_)
    t = np.linspace(0.0, 5e-3, 400)
    B, C, _ = getattr(orc, method)(f, t, mask, L=L, n_bins=4096)
    if L > 0:
        assert B.shape == (400, L) and C.shape == (L, *f.shape)
    E_full = np.exp(np.outer(t, f[mask]))
    np.testing.assert_allclose(B @ C[:, mask], E_full, atol=5e-3, rtol=5e-3)


def test_histogram_counts_and_centres():
    f = np.array([[1j, 2j], [3j, 3j]])  # constant real part: (1, n_bins) bins
    h, w = orc.histogram(f, None, 4)
    assert h.shape == (1, 4) and h.sum() == 4
    np.testing.assert_array_equal(h.ravel(), [1, 0, 1, 2])
    np.testing.assert_allclose(w.ravel().imag, [1.25, 1.75, 2.25, 2.75])


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("method", ["svd", "mti"])
def test_host_factorizations_match_the_oracle(method, complex_):
    """hipnufft.field_map (torch, complex128) against the numpy restatement:
    the same histogram and bin mapping, the same product B C."""
    import torch

    from hipnufft import field_map as hfm

    f, mask = _maps(32, complex_)
    t = np.linspace(0.0, 5e-3, 300).astype(np.float32)
    fc = f.astype(np.complex64)
    hh, hw = hfm._create_histogram(torch.from_numpy(fc), torch.from_numpy(mask), 2048)
    oh, ow = orc.histogram(fc, mask, 2048)
    assert hh.shape == oh.shape
    assert np.abs(hh.numpy() - oh).sum() <= 2  # edge-case rounding of bin edges
    np.testing.assert_allclose(hw.numpy(), ow, rtol=1e-6, atol=1e-3)
    B, C, _ = hfm.get_orc_factorization(method)(fc, t, mask, L=20, n_bins=2048)
    rB, rC, _ = getattr(orc, method)(fc, t, mask, L=20, n_bins=2048)
    got = B @ C.reshape(20, -1)
    ref = rB @ rC.reshape(20, -1)
    assert rel_l2(got, ref) < 1e-4


def test_host_mfi_and_lazy():
    from hipnufft import field_map as hfm

    f, mask = _maps(32, False)
    t = np.linspace(0.0, 5e-3, 400).astype(np.float32)
    B, C, _ = hfm.get_orc_factorization("mfi")(f.astype(np.complex64), t, mask, L=40,
                                               n_bins=4096)
    assert len(C) == 40 and C.shape == (40, 32, 32)
    Cf = np.stack([C[l] for l in range(len(C))])
    np.testing.assert_allclose(B @ Cf[:, mask], np.exp(np.outer(t, f[mask])), atol=5e-3,
                               rtol=5e-3)


def test_make_b0map_properties():
    from hipnufft import make_b0map

    b0, mask = make_b0map((40, 40), b0range=(-200, 100))
    assert b0.dtype == np.float32 and mask.dtype == bool
    assert np.all(b0[~mask] == 0)
    assert b0[mask].min() >= -200 - 1e-3 and np.isclose(b0[mask].max(), 100)
    assert b0[20, 20] == pytest.approx(-200)  # the centre has radius 0
