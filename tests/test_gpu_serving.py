"""Continuous-batching server: results alone vs in the mix, admission by HBM
budget, overlap of requests, and agreement with the plain per-call CG."""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, radial2d, rand_cplx, rand_traj, rel_l2

pytestmark = pytest.mark.gpu


def _requests():
    from hipnufft.serving import ReconRequest, synth_kspace

    specs = [  # (key, traj, shape, C)
        ("radA", radial2d(24, 128), (48, 48), 3),
        ("radA", radial2d(24, 128), (48, 48), 3),
        ("radB", radial2d(40, 96), (48, 48), 1),
        ("rnd", rand_traj(4000, 2, 61), (40, 56), 4),
        ("vol", rand_traj(6000, 3, 62), (16, 16, 12), 2),
        ("radA", radial2d(24, 128), (48, 48), 2),
    ]
    reqs = []
    for i, (key, traj, shape, C) in enumerate(specs):
        smaps = gauss_smaps(C, shape, seed=70 + i)
        img = rand_cplx(shape, seed=80 + i)
        y, _ = synth_kspace(traj, shape, C, smaps, img)
        reqs.append(ReconRequest(traj, shape, y, smaps, max_iter=6, traj_key=key,
                                 arrival=0.002 * i, rid=i))
    return reqs


@pytest.mark.parametrize("toeplitz", [True, False])
def test_alone_vs_mix_identical(toeplitz):
    from hipnufft.serving import ReconServer

    reqs = _requests()
    mix = ReconServer(max_slots=4, toeplitz=toeplitz).run(reqs, realtime=False)
    assert sorted(mix) == list(range(len(reqs)))
    # all requests wait from the start: latencies count from the call
    assert all(m.arrival == 0.0 and 0.0 <= m.admitted <= m.finished for m in mix.values())
    # every iteration enqueued at admission: the same solves
    deep = ReconServer(max_slots=4, toeplitz=toeplitz, depth=0).run(reqs, realtime=False)
    for r in reqs:
        assert rel_l2(deep[r.rid].image, mix[r.rid].image) <= 1e-6, r.rid
        alone = ReconServer(max_slots=1, toeplitz=toeplitz).run([r], realtime=False)[r.rid]
        assert rel_l2(mix[r.rid].image, alone.image) <= 1e-6, r.rid


@pytest.mark.parametrize("toeplitz,tol", [(False, 1e-5), (True, 1e-4)])
def test_served_matches_plain_cg(toeplitz, tol):
    """Served results against the plain per-call CG with the same L (the
    Toeplitz gradients differ from the NUFFT ones at the NUFFT tolerance;
    the stated reconstruction bar is 1e-4)."""
    from hipnufft import MRIHipNUFFT, cg
    from hipnufft.serving import ReconServer

    reqs = _requests()[:4]
    srv = ReconServer(max_slots=4, toeplitz=toeplitz)
    res = srv.run(reqs, realtime=False)
    for r in reqs:
        op = MRIHipNUFFT(r.samples, r.shape, n_coils=r.kspace.shape[0], smaps=r.smaps,
                         squeeze_dims=False)
        L = next(p.lipschitz for k, p in srv.plans.items()
                 if k[0] == r.traj_key or (isinstance(k[0], tuple) and k[0][0] == r.traj_key))
        op.get_lipschitz_cst = lambda **_: L
        y = torch.from_numpy(r.kspace).to(op.device).reshape(1, r.kspace.shape[0], -1)
        ref = cg(op, y, max_iter=r.max_iter, tol=r.tol).reshape(r.shape).cpu().numpy()
        assert rel_l2(res[r.rid].image, ref) <= tol, r.rid


def test_budget_admission_and_reuse():
    from hipnufft.serving import ReconServer, estimate_bytes, plan_ntrans

    reqs = _requests()
    big = max(estimate_bytes(r.shape, r.kspace.shape[1],
                             max(r.kspace.shape[0], plan_ntrans(r.shape, r.kspace.shape[0])),
                             toeplitz=True) for r in reqs)
    srv = ReconServer(hbm_budget_bytes=int(1.5 * big), max_slots=4)
    res = srv.run(reqs, realtime=False)
    free = ReconServer(max_slots=4).run(reqs, realtime=False)
    for r in reqs:
        assert rel_l2(res[r.rid].image, free[r.rid].image) <= 1e-6
    assert srv.stats["evictions"] > 0
    assert srv.stats["admitted"] == len(reqs)
    reuse = ReconServer(max_slots=1)
    reuse.run([reqs[0], reqs[1]], realtime=False)
    assert reuse.stats["plan_hits"] == 1 and reuse.stats["plan_builds"] == 1


def test_requests_overlap_on_device():
    """A request that arrives while another is being solved is admitted
    before the first finishes (continuous batching, not a queue)."""
    from hipnufft.serving import ReconServer

    reqs = _requests()
    first, second = reqs[4], reqs[0]   # a 3D solve, then a small 2D one
    first.arrival, second.arrival = 0.0, 1e-4
    first.max_iter = 30
    res = ReconServer(max_slots=2).run([first, second], realtime=True)
    assert res[second.rid].admitted < res[first.rid].finished
