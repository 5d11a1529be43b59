"""Pins of the FP64 NDFT oracle (CPU).

Mirrors the reference's own NDFT tests:
* known answer: Cartesian fftfreq samples -> fftn(fftshift(img))
  (tests/test_ndft.py:58-79, cases grid1D / grid2D / grid3D of
  tests/case_trajectories.py:64-82);
* explicit matrix == implicit loops (tests/test_ndft.py:19-55);
plus adjointness and a scalar brute-force sum on a tiny input.
"""

import numpy as np
import pytest
import scipy as sp

from helpers import rand_cplx, truncnorm_traj
from oracle import ndft


def grid_case(N, d):
    f = sp.fft.fftfreq(N)
    g = np.stack(np.meshgrid(*([f] * d)), axis=-1)
    return g.reshape(-1, d), (N,) * d


@pytest.mark.parametrize("N,d", [(256, 1), (16, 2), (16, 3)])
def test_ndft_grid_known_answer(N, d):
    k, shape = grid_case(N, d)
    img = rand_cplx(shape, 1).astype(np.complex128)
    ks = ndft.type2(k, img, shape).reshape(shape)
    if d >= 2:
        ks = ks.swapaxes(0, 1)  # meshgrid "xy" indexing of the sample grid
    np.testing.assert_allclose(ks, sp.fft.fftn(sp.fft.fftshift(img)), rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("case", ["random2D", "grid2D", "grid3D"])
def test_ndft_explicit_vs_implicit(case):
    if case == "random2D":
        k, shape = truncnorm_traj(1000, 2, 0), (64, 128)
    else:
        k, shape = grid_case(16, 2 if case == "grid2D" else 3)
    mat = ndft.fourier_matrix(k, shape)
    x = rand_cplx(shape, 2)
    y = rand_cplx(len(k), 3)
    np.testing.assert_allclose(ndft.type2(k, x, shape), mat @ x.ravel(), rtol=1e-9, atol=1e-8)
    np.testing.assert_allclose(ndft.type1(k, y, shape).ravel(), mat.conj().T @ y,
                               rtol=1e-9, atol=1e-8)


def test_ndft_adjoint_and_norm():
    k = truncnorm_traj(300, 2, 4)
    shape = (12, 10)
    x = rand_cplx(shape, 5).astype(np.complex128)
    y = rand_cplx(len(k), 6).astype(np.complex128)
    lhs = np.vdot(ndft.type2(k, x, shape, normalize=True), y)
    rhs = np.vdot(x, ndft.type1(k, y, shape, normalize=True))
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)
    assert ndft.norm_factor((12, 10)) == pytest.approx(np.sqrt(120 * 4))


def test_ndft_brute_scalar():
    k = np.array([[0.1, -0.3], [0.45, 0.2], [-0.5, 0.0]])
    shape = (4, 2)
    x = rand_cplx(shape, 7).astype(np.complex128)
    out = np.zeros(len(k), dtype=complex)
    for m in range(len(k)):
        for i in range(4):
            for j in range(2):
                out[m] += x[i, j] * np.exp(-2j * np.pi * (k[m, 0] * (i - 2) + k[m, 1] * (j - 1)))
    np.testing.assert_allclose(ndft.type2(k, x, shape), out, rtol=1e-12)


def test_unit_trajectory_rescale():
    rad = np.array([[np.pi / 2, -np.pi]])
    np.testing.assert_allclose(ndft.unit_trajectory(rad), [[0.25, -0.5]])
    unit = np.array([[0.25, 0.5]])
    np.testing.assert_allclose(ndft.unit_trajectory(unit), [[0.25, -0.5]])


@pytest.mark.parametrize("shape", [(16,), (12, 10), (8, 6, 10)])
def test_ndft_gpu_matches_oracle(shape):
    """The separable FP64 NDFT used by the large-size GPU tests
    (``tests/ndft_gpu.py``, run here on torch CPU tensors) equals the
    oracle's explicit sums, full and at sample / voxel subsets."""
    import torch

    import ndft_gpu

    d = len(shape)
    k = truncnorm_traj(300, d, 11).astype(np.float64)
    om = torch.from_numpy(2 * np.pi * k)
    x = rand_cplx((2, *shape), 12).astype(np.complex128)
    y = rand_cplx((2, len(k)), 13).astype(np.complex128)
    want2 = np.stack([ndft.type2(k, x[t], shape) for t in range(2)])
    want1 = np.stack([ndft.type1(k, y[t], shape) for t in range(2)])
    got2 = ndft_gpu.type2(om, torch.from_numpy(x), shape, chunk=77).numpy()
    got1 = ndft_gpu.type1(om, torch.from_numpy(y), shape, chunk=77).numpy()
    np.testing.assert_allclose(got2, want2, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(got1, want1, rtol=1e-12, atol=1e-10)
    idx = torch.tensor([0, 5, 299, 17])
    np.testing.assert_allclose(ndft_gpu.type2_at(om, torch.from_numpy(x), shape, idx).numpy(),
                               want2[:, idx.numpy()], rtol=1e-12, atol=1e-10)
    vox = torch.tensor([0, 3, int(np.prod(shape)) - 1])
    np.testing.assert_allclose(ndft_gpu.type1_at(om, torch.from_numpy(y), shape, vox, chunk=50).numpy(),
                               want1.reshape(2, -1)[:, vox.numpy()], rtol=1e-12, atol=1e-10)
