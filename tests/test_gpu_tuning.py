"""Autotuner: candidates run, the winner is cached, and a tuned operator
gives the same results as the default one (to FP32 summation-order noise)."""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, radial2d, rand_cplx, rel_l2

pytestmark = pytest.mark.gpu


def test_autotune_runs_and_caches():
    from hipnufft import autotune as at

    traj = radial2d(64, 256)
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).cuda()
    res = at.autotune(om, (96, 96), n_trans=4, reps=1)
    assert res["_tried"] >= 5 and res["_ms"] > 0
    opts = at.plan_options(res)
    assert set(opts) == {"tile", "spread_group", "interp_group", "max_subproblem", "spread_max",
                         "spread_method"}
    again = at.autotune(om, (96, 96), n_trans=4, reps=1)
    assert again == res


def test_tuned_operator_matches_default():
    from hipnufft import MRIHipNUFFT

    traj = radial2d(48, 200)
    shape = (80, 72)
    smaps = gauss_smaps(4, shape, seed=3)
    a = MRIHipNUFFT(traj, shape, n_coils=4, smaps=smaps, autotune=True)
    b = MRIHipNUFFT(traj, shape, n_coils=4, smaps=smaps)
    assert a.tuning is not None
    x = torch.from_numpy(rand_cplx(shape, seed=4)).cuda()
    ya, yb = a.op(x), b.op(x)
    assert rel_l2(ya.cpu().numpy(), yb.cpu().numpy()) < 1e-5
    assert rel_l2(a.adj_op(yb).cpu().numpy(), b.adj_op(yb).cpu().numpy()) < 1e-5


def test_explicit_option_kept_over_tuning():
    """An option the caller sets is not replaced by the tuner's choice."""
    from hipnufft import MRIHipNUFFT

    traj = radial2d(32, 128)
    op = MRIHipNUFFT(traj, (48, 48), n_coils=2, autotune=True, spread_group=2,
                     max_subproblem=333)
    assert op.tuning is not None
    assert op._plan_opts["spread_group"] == 2 and op._plan_opts["max_subproblem"] == 333
