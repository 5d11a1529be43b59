"""LSQR and LSMR of the GPU operator against scipy's published solvers on the
exact FP64 NDFT operator (``oracle/solvers.py``: ``lsmr``, ``sense_linop``;
``oracle/smaps.py``: ``lsqr``), which the reference's ``lsqr`` / ``lsmr``
(``extras/optim.py:249-798``) adapt.  The operator's float32 radians are
handed to the oracle, so both solve the same problem.  SENSE and
calibrationless, damped and undamped, one batch and several (each batch is
an independent problem for scipy).
"""

import numpy as np
import pytest

from helpers import gauss_smaps, rand_cplx, rel_l2, truncnorm_traj
from oracle import solvers

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _scipy_solve(optim, A, b, it, damp):
    if optim == "lsmr":
        return solvers.lsmr(A, b, it, damp=damp)
    from scipy.sparse.linalg import lsqr

    return lsqr(A, b, damp=damp, atol=1e-6, btol=1e-6, conlim=1e8, iter_lim=it)[0]


@pytest.mark.parametrize("optim", ["lsmr", "lsqr"])
@pytest.mark.parametrize("sense,damp", [(True, 0.0), (False, 0.0), (True, 0.5)])
def test_solver_vs_scipy_on_exact_ndft(optim, sense, damp):
    from hipnufft import MRIHipNUFFT

    shape, C, it = (24, 20), 3, 10
    traj = truncnorm_traj(700, 2, 41)
    smaps = gauss_smaps(C, shape, 42) if sense else None
    op = MRIHipNUFFT(traj, shape, n_coils=C, smaps=smaps, squeeze_dims=True)
    y = rand_cplx((C, len(traj)), 43)
    x = op.pinv_solver(y, optim=optim, max_iter=it, damp=damp)
    A = solvers.sense_linop(op.samples.cpu().numpy(), shape, smaps=smaps, n_coils=C)
    ref = _scipy_solve(optim, A, y.reshape(-1).astype(np.complex128), it, damp)
    err = rel_l2(np.asarray(x).reshape(-1), ref)
    print(f"{optim} sense={sense} damp={damp}: rel L2 vs scipy {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("optim", ["lsmr", "lsqr"])
def test_batched_solver_vs_scipy_per_batch(optim):
    """n_batchs = 3: per-batch scalars; each batch equals scipy on its own data."""
    from hipnufft import MRIHipNUFFT

    shape, C, B, it = (20, 16), 2, 3, 8
    traj = truncnorm_traj(500, 2, 44)
    smaps = gauss_smaps(C, shape, 45)
    op = MRIHipNUFFT(traj, shape, n_coils=C, n_batchs=B, smaps=smaps)
    y = rand_cplx((B, C, len(traj)), 46)
    y[1] *= 10.0                                  # batches of different scale
    x = np.asarray(op.pinv_solver(y, optim=optim, max_iter=it)).reshape(B, -1)
    A = solvers.sense_linop(op.samples.cpu().numpy(), shape, smaps=smaps)
    for b in range(B):
        ref = _scipy_solve(optim, A, y[b].reshape(-1).astype(np.complex128), it, 0.0)
        assert rel_l2(x[b], ref) < 1e-4, b

