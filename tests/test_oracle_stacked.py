"""Pins of the stacked 2.5D oracle (CPU): for an even number of z planes
the stacked transform equals the 3D NDFT of the equivalent 3D trajectory
(planes at kz = (z - Z//2) / Z), forward and adjoint; the oracle is its own
adjoint's adjoint; the trajectory conversions round-trip
(tests/operators/test_stacked.py:121-133 of the reference)."""

import numpy as np
import pytest

from helpers import rand_cplx, truncnorm_traj
from oracle import ndft, stacked


def _case(sense, B=2, C=3, X=8, Y=10, Z=6, NS=40, seed=0):
    rng = np.random.default_rng(seed)
    s2 = truncnorm_traj(NS, 2, seed)
    zi = rng.choice(Z, 4, replace=False)
    shape = (X, Y, Z)
    smaps = rand_cplx((C, *shape), seed + 1).astype(np.complex128) if sense else None
    img = rand_cplx((B, *shape) if sense else (B, C, *shape), seed + 2).astype(np.complex128)
    return s2, zi, shape, smaps, img


def _traj3d(s2, zi, Z):
    z = (np.asarray(zi) - Z // 2) / Z
    out = np.zeros((len(zi), len(s2), 3))
    out[:, :, :2] = s2[None]
    out[:, :, 2] = z[:, None]
    return out.reshape(-1, 3)


@pytest.mark.parametrize("sense", [False, True])
def test_stacked_equals_3d_ndft(sense):
    s2, zi, shape, smaps, img = _case(sense)
    B, C = 2, 3
    y = stacked.op(s2, zi, img, shape, smaps)
    k3 = _traj3d(s2, zi, shape[2])
    coil = img[:, None] * smaps[None] if sense else img
    ref = np.stack([np.stack([ndft.type2(k3, coil[b, c], shape) for c in range(C)])
                    for b in range(B)]) / ndft.norm_factor(shape)
    np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-12)
    rng = np.random.default_rng(5)
    ksp = rand_cplx(y.shape, 6)
    dens = rng.uniform(0.5, 2.0, len(s2))
    x = stacked.adj(s2, zi, ksp, shape, smaps, density=dens, n_batchs=B)
    d3 = np.tile(dens, len(zi))
    imgs = np.stack([np.stack([ndft.type1(k3, ksp[b, c] * d3, shape) for c in range(C)])
                     for b in range(B)]) / ndft.norm_factor(shape)
    if sense:
        imgs = np.sum(imgs * np.conj(smaps)[None], axis=1, keepdims=True)
    np.testing.assert_allclose(x, imgs, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("Z", [5, 7, 6])
def test_stacked_oracle_adjoint(Z):
    s2, zi, shape, smaps, img = _case(True, Z=Z, seed=3)
    y = stacked.op(s2, zi, img, shape, smaps)
    ksp = rand_cplx(y.shape, 8)
    x = stacked.adj(s2, zi, ksp, shape, smaps, n_batchs=2)
    assert abs(np.vdot(y, ksp) - np.vdot(img.reshape(x.shape), x)) < 1e-10 * abs(np.vdot(y, ksp))


def test_stack_trajectory_round_trip():
    from hipnufft.stacked import stacked2traj3d, traj3d2stacked

    dimz = 64
    rng = np.random.default_rng(0)
    traj2d = rng.uniform(-0.5, 0.5, (100, 2)).astype(np.float32)
    z_index = rng.choice(dimz, 20, replace=False)
    traj3d = stacked2traj3d(traj2d, z_index, dimz)
    t2, z2 = traj3d2stacked(traj3d, dimz)
    np.testing.assert_allclose(t2, traj2d)
    np.testing.assert_array_equal(z2, z_index)
