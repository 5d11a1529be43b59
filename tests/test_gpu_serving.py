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


@pytest.mark.parametrize("toeplitz,graph", [(True, False), (False, False), (True, True)])
def test_alone_vs_mix_identical(toeplitz, graph):
    from hipnufft.serving import ReconServer

    reqs = _requests()
    mix = ReconServer(max_slots=4, toeplitz=toeplitz, use_graph=graph).run(reqs, realtime=False)
    assert sorted(mix) == list(range(len(reqs)))
    # all requests wait from the start: latencies count from the call
    assert all(m.arrival == 0.0 and 0.0 <= m.admitted <= m.finished for m in mix.values())
    # every iteration enqueued at admission: the same solves
    deep = ReconServer(max_slots=4, toeplitz=toeplitz, depth=0).run(reqs, realtime=False)
    for r in reqs:
        assert rel_l2(deep[r.rid].image, mix[r.rid].image) <= 1e-6, r.rid
        alone = ReconServer(max_slots=1, toeplitz=toeplitz).run([r], realtime=False)[r.rid]
        assert rel_l2(mix[r.rid].image, alone.image) <= 1e-6, r.rid  # graph replays == eager


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


def test_invalid_requests_do_not_disturb_others():
    """Malformed requests (coil maps for the wrong coil count, k-space of the
    wrong length, over the HBM budget) placed between good ones are answered
    with an error; the good results are bit-identical to their alone runs,
    and the server stays usable."""
    from hipnufft.serving import ReconRequest, ReconServer

    good = _requests()[:3]
    bad_maps = ReconRequest(good[0].samples, good[0].shape, good[0].kspace,
                            good[0].smaps[:2], max_iter=6, traj_key="radA", rid=10)
    bad_ksp = ReconRequest(good[2].samples, good[2].shape, good[2].kspace[:, :-5],
                           good[2].smaps, max_iter=6, rid=11)
    huge_shape = (2048, 2048, 64)
    huge = ReconRequest(rand_traj(100, 3, 5), huge_shape, rand_cplx((1, 100), 6),
                        np.ones((1, 1, 1, 1), np.complex64), max_iter=3, rid=12)
    huge.smaps = np.broadcast_to(np.ones((1,) + huge_shape, np.complex64), (1,) + huge_shape)
    mix = [good[0], bad_maps, good[1], bad_ksp, huge, good[2]]
    srv = ReconServer(hbm_budget_bytes=8 << 30, max_slots=4)
    res = srv.run(mix, realtime=False)
    assert sorted(res) == sorted(r.rid for r in mix)
    for r in (bad_maps, bad_ksp, huge):
        assert not res[r.rid].ok and res[r.rid].image is None
    assert "coil maps" in res[bad_maps.rid].error
    assert "k-space" in res[bad_ksp.rid].error
    assert "budget" in res[huge.rid].error
    assert srv.stats["failed"] == 3
    for r in good:
        alone = ReconServer(max_slots=1).run([r], realtime=False)[r.rid]
        assert res[r.rid].ok
        np.testing.assert_array_equal(res[r.rid].image, alone.image)
    # the same server serves again, with every plan idle
    assert all(not p.in_use for p in srv.plans.values())
    again = srv.run([good[1]], realtime=False)
    np.testing.assert_array_equal(again[good[1].rid].image, res[good[1].rid].image)


def test_complex128_maps_are_cast():
    """Coil maps given as complex128 give the complex64 maps' result (the
    server casts them instead of reading their bytes as complex64)."""
    import copy

    from hipnufft.serving import ReconServer

    r64 = _requests()[3]
    r128 = copy.copy(r64)
    r128.smaps = r64.smaps.astype(np.complex128)
    a = ReconServer(max_slots=1).run([r64], realtime=False)[r64.rid].image
    b = ReconServer(max_slots=1).run([r128], realtime=False)[r128.rid].image
    np.testing.assert_array_equal(a, b)


def _serving_kinds():
    """One request of each kind of the serving workload at its real size:
    2D radial 256^2 (256 x 512 samples, 32 coils), 2D spiral 384^2
    (96 x 4096, 16 coils), 3D stack of stars 128^3 (256 x 256 on 128
    Cartesian partitions, 8 coils); 10 CG iterations each."""
    from hipnufft import stacked2traj3d
    from hipnufft.serving import ReconRequest, synth_kspace
    from hipnufft.trajectories import gaussian_smaps, initialize_2D_radial, initialize_2D_spiral

    rng = np.random.default_rng(55)
    stars = stacked2traj3d(initialize_2D_radial(256, 256).reshape(-1, 2), np.arange(128), 128)
    specs = [("radial2d", initialize_2D_radial(256, 512).reshape(-1, 2), (256, 256), 32),
             ("spiral2d", initialize_2D_spiral(96, 4096).reshape(-1, 2), (384, 384), 16),
             ("stars3d", stars, (128, 128, 128), 8)]
    reqs = []
    for i, (key, traj, shape, C) in enumerate(specs):
        traj = traj.astype(np.float32)
        smaps = gaussian_smaps(C, shape, seed=int(rng.integers(1 << 30)))
        img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        y, _ = synth_kspace(traj, shape, C, smaps, img)
        reqs.append(ReconRequest(traj, shape, y, smaps, max_iter=10, traj_key=(key,),
                                 arrival=0.0, rid=i))
    return reqs


def test_serving_kinds_alone_mix_and_plain():
    """Each kind alone and in the mix gives the same image; the served
    result (resident plans, Toeplitz CG; the stack of stars on the stacked
    operator) equals the plain per-call CG on a fresh full 3D / 2D operator
    to the reconstruction bar."""
    from hipnufft.serving import ReconServer, serve_plain

    reqs = _serving_kinds()
    srv = ReconServer(max_slots=8)
    mix = srv.run(reqs, realtime=False)
    assert [p.stacked for k, p in srv.plans.items() if k[0][0] == "stars3d"] == [True]
    plain = serve_plain(reqs, realtime=False)
    for r in reqs:
        alone = ReconServer(max_slots=1).run([r], realtime=False)[r.rid]
        e_mix = rel_l2(mix[r.rid].image, alone.image)
        e_plain = rel_l2(mix[r.rid].image, plain[r.rid].image)
        print(f"{r.traj_key[0]}: mix vs alone {e_mix:.1e}, served vs plain {e_plain:.1e}")
        assert e_mix <= 1e-6
        assert e_plain <= 1e-4
