"""Shared test helpers (seeded inputs, error metrics)."""

import numpy as np


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rand_traj(M, d, seed=0):
    """Uniform samples in [-0.5, 0.5) (float32)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.5, 0.5, (M, d)).astype(np.float32)


def truncnorm_traj(M, d, seed=0):
    """Gaussian-like samples concentrated at the centre (like case_random2D)."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 0.16, (M * 4, d))
    x = x[np.all(np.abs(x) < 0.48, axis=1)][:M]
    return x.astype(np.float32)


def radial2d(Nc, Ns, in_out=False):
    """2D radial spokes (closed form of initialize_2D_radial)."""
    r = 0.5 * np.linspace(-1 if in_out else 0, 1, Ns)
    th = 2 * np.pi * np.arange(Nc) / Nc / (1 + in_out)
    t = np.stack([np.cos(th)[:, None] * r, np.sin(th)[:, None] * r], -1)
    return t.reshape(-1, 2).astype(np.float32)


def rand_cplx(shape, seed=0, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def gauss_smaps(C, shape, seed=0, sigma=0.3):
    """Coil maps: sums of 4 complex Gaussian bumps, normalised to unit RSS."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.linspace(-0.5, 0.5, s, endpoint=False) for s in shape],
                        indexing="ij")
    maps = np.zeros((C, *shape), dtype=np.complex128)
    for c in range(C):
        for _ in range(4):
            ctr = rng.uniform(-0.5, 0.5, len(shape))
            amp = rng.standard_normal() + 1j * rng.standard_normal()
            r2 = sum((g - x) ** 2 for g, x in zip(grids, ctr))
            maps[c] += amp * np.exp(-r2 / (2 * sigma ** 2))
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    return (maps / rss).astype(np.complex64)


def assert_correlate(a, b, slope=1.0, slope_err=1e-3, r_value_err=1e-3):
    """Correlation bar of the reference's tests (``tests/helpers/asserts.py:50-70``):
    over real and imaginary parts, |r - 1| <= r_value_err and the regression
    slope of b on a within slope_err of ``slope``; a bar given as None is
    not checked, as in the reference."""
    a = np.concatenate([np.real(a).ravel(), np.imag(a).ravel()]).astype(np.float64)
    b = np.concatenate([np.real(b).ravel(), np.imag(b).ravel()]).astype(np.float64)
    r = np.corrcoef(a, b)[0, 1]
    s = abs(np.cov(a, b)[0, 1] / np.var(a))
    if r_value_err is not None:
        assert abs(r - 1) <= r_value_err, f"r = {r}"
    if slope_err is not None:
        assert abs(s - slope) <= slope_err, f"slope = {s}"
