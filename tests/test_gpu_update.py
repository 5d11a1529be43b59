"""Updates of samples, density and coil maps against the oracle.

Mirrors the reference's ``tests/operators/test_update.py:131-248`` (jittered
samples through the ``samples`` setter, density jitter / added later /
removed, new coil maps; op and adj_op after each), but the expected values
come from the exact FP64 operator of the oracle (``oracle/mri.py``) built on
the updated inputs, not from a second hipnufft operator.  Also checks that
``update_samples`` re-runs the density method (Pipe against the oracle's
Pipe, ``cufinufft.py:317-347``) and rejects a failed update without
leaving the operator on stale samples.
"""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, rand_cplx, rel_l2, truncnorm_traj
from oracle import mri, solvers

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (n_batchs, n_coils, sense) of the reference's operator fixture
LAYOUTS = [(1, 1, False), (3, 1, False), (1, 4, True), (1, 4, False), (3, 2, True),
           (3, 4, False)]


def _case(dim):
    if dim == 2:
        return truncnorm_traj(1000, 2, 60), (64, 48)
    return truncnorm_traj(3000, 3, 61), (16, 20, 12)


def _op(traj, shape, B, C, sense, density=False):
    from hipnufft import MRIHipNUFFT

    smaps = gauss_smaps(C, shape, 62) if sense else None
    return MRIHipNUFFT(traj, shape, n_coils=C, n_batchs=B, smaps=smaps, density=density), smaps


def _oracle(op, smaps=None, density=None):
    """Exact operator on the operator's current float32 radians."""
    om = op.samples.detach().cpu().numpy()
    sm = None if smaps is None else np.asarray(smaps)
    d = None if density is None else np.asarray(density.detach().cpu() if
                                                 isinstance(density, torch.Tensor) else density)
    return mri.Operator(om, op.shape, n_coils=op.n_coils, n_batchs=op.n_batchs, smaps=sm,
                        density=d, exact=True)


def _x(op, seed):
    ic = 1 if op.uses_sense else op.n_coils
    return rand_cplx((op.n_batchs, ic, *op.shape), seed)


def _y(op, seed):
    return rand_cplx((op.n_batchs, op.n_coils, op.n_samples), seed)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("B,C,sense", LAYOUTS)
def test_samples_setter_jitter(dim, B, C, sense):
    traj, shape = _case(dim)
    op, smaps = _op(traj, shape, B, C, sense)
    jitter = np.random.default_rng(63).random(op.samples.shape).astype(np.float32)
    op.samples += torch.from_numpy(jitter / 100).to(DEV)     # the reference's in-place update
    ref = _oracle(op, smaps)
    x, y = _x(op, 64), _y(op, 65)
    assert rel_l2(op.op(x), ref.op(x)) < 2e-5
    assert rel_l2(op.adj_op(y), ref.adj_op(y)) < 2e-5


@pytest.mark.parametrize("B,C,sense", LAYOUTS)
def test_density_jitter_added_removed(B, C, sense):
    traj, shape = _case(2)
    rng = np.random.default_rng(66)
    d0 = rng.uniform(0.5, 1.5, len(traj)).astype(np.float32)
    op, smaps = _op(traj, shape, B, C, sense, density=d0)
    y = _y(op, 67)
    # jitter in place through the setter
    op.density += torch.from_numpy(rng.random(len(traj)).astype(np.float32) / 100).to(DEV)
    assert rel_l2(op.adj_op(y), _oracle(op, smaps, op.density).adj_op(y)) < 2e-5
    # removed
    op.density = None
    assert rel_l2(op.adj_op(y), _oracle(op, smaps).adj_op(y)) < 2e-5
    # added later (a numpy array)
    d1 = (1e-2 + rng.random(len(traj)) / 100).astype(np.float32)
    op.density = d1
    assert rel_l2(op.adj_op(y), _oracle(op, smaps, d1).adj_op(y)) < 2e-5
    # data consistency applies the new density once
    x = _x(op, 68)
    assert rel_l2(op.data_consistency(x, y), _oracle(op, smaps, d1).data_consistency(x, y)) < 2e-5


@pytest.mark.parametrize("B,C", [(1, 4), (3, 2)])
def test_smaps_update(B, C):
    traj, shape = _case(2)
    op, _ = _op(traj, shape, B, C, True)
    rng = np.random.default_rng(69)
    new = (rng.random((C, *shape)) + 1j * rng.random((C, *shape))).astype(np.complex64)
    new /= np.linalg.norm(new, axis=0)
    op.smaps = new
    ref = _oracle(op, new)
    x, y = _x(op, 70), _y(op, 71)
    assert rel_l2(op.op(x), ref.op(x)) < 2e-5
    assert rel_l2(op.adj_op(y), ref.adj_op(y)) < 2e-5
    # complex128 maps are cast, not misread
    op.smaps = new.astype(np.complex128)
    assert rel_l2(op.op(x), ref.op(x)) < 2e-5


def test_update_samples_recomputes_pipe():
    """``density=True`` stores Pipe as the density method; ``update_samples``
    re-runs it on the new samples (``cufinufft.py:317-347``)."""
    from hipnufft.trajectories import initialize_2D_radial

    shape = (32, 32)
    t1 = initialize_2D_radial(128, 16).reshape(-1, 2).astype(np.float32)
    t2 = initialize_2D_radial(128, 16, tilt="golden").reshape(-1, 2).astype(np.float32)
    op, _ = _op(t1, shape, 1, 1, False, density=True)
    d1 = op.density.cpu().numpy()
    assert rel_l2(d1, solvers.pipe(t1, shape, max_iter=10)) < 1e-3
    op.update_samples(t2)
    d2 = op.density.cpu().numpy()
    assert rel_l2(d2, solvers.pipe(t2, shape, max_iter=10)) < 1e-3
    assert rel_l2(d2, d1) > 1e-2
    y = _y(op, 72)
    assert rel_l2(op.adj_op(y), _oracle(op, None, d2).adj_op(y)) < 2e-5


def test_failed_update_keeps_operator_usable():
    """An update with the wrong dimension raises and leaves the operator on
    its previous samples."""
    traj, shape = _case(2)
    op, _ = _op(traj, shape, 1, 2, False)
    x = _x(op, 73)
    before = op.op(x)
    with pytest.raises(ValueError):
        op.update_samples(np.zeros((10, 3), np.float32))
    assert op.n_samples == len(traj)
    np.testing.assert_array_equal(op.op(x), before)


def test_failed_setpts_leaves_plan_unusable_until_rebuilt(monkeypatch):
    """A device sort that fails part-way (forced from the host through
    HIPNUFFT_INJECT_FAILURE, after the points are permuted and before the
    spread lists are built) leaves a plan that refuses to transform (no
    stale indices are used), the operator on its previous samples, and a
    later successful update restores it."""
    from hipnufft._capi import HipNufftError

    traj, shape = _case(2)
    op, _ = _op(traj, shape, 1, 2, False)
    x = _x(op, 74)
    good = op.op(x)
    monkeypatch.setenv("HIPNUFFT_INJECT_FAILURE", "setpts")
    with pytest.raises(HipNufftError, match="injected"):
        op.update_samples(truncnorm_traj(1500, 2, 75))
    monkeypatch.delenv("HIPNUFFT_INJECT_FAILURE")
    with pytest.raises(ValueError, match="setpts"):
        op.op(x)
    with pytest.raises(ValueError, match="setpts"):
        op.adj_op(_y(op, 76))
    assert op.n_samples == len(traj)
    op.update_samples(traj)
    np.testing.assert_array_equal(op.op(x), good)
