"""Several trajectories of uneven length in one plan (point sets).

* the joint bin sort is bit-exact against the oracle with set-major keys
  (empty sets and very uneven lengths included);
* op / adj_op / data consistency of the multi-trajectory operator equal
  those of one MRIHipNUFFT per trajectory (same kernels, same summation
  order: the bar is 1e-6 relative, in practice bit-identical), whatever the
  chunking (one set per chunk or all sets at once);
* the batched CG equals per-problem CG with the same step size;
* row-wise FP64 dots equal single dots bit for bit.
"""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, radial2d, rand_cplx, rand_traj, rel_l2, truncnorm_traj
from oracle import nufft

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _trajs_2d():
    return [rand_traj(900, 2, 21), radial2d(6, 77), np.zeros((0, 2), np.float32),
            truncnorm_traj(2500, 2, 22), rand_traj(3, 2, 23)]


def _trajs_3d():
    return [rand_traj(1500, 3, 31), rand_traj(400, 3, 32) * 0.1, rand_traj(2600, 3, 33)]


@pytest.mark.parametrize("dim", [2, 3])
def test_sets_sort_bitexact(dim):
    from hipnufft import _capi

    trajs = _trajs_2d() if dim == 2 else _trajs_3d()
    shape = (40, 36) if dim == 2 else (16, 20, 12)
    sizes = [t.shape[0] for t in trajs]
    om = [(t * np.float32(2 * np.pi)).astype(np.float32) for t in trajs]
    p = _capi.RawPlan(shape, n_trans=1)
    p.setpts_sets([torch.from_numpy(o).to(DEV) for o in om])
    info = p.info()
    nf = info["nf"][3 - dim:]
    tile = info["tile"][3 - dim:]
    allom = np.concatenate(om, 0)
    perm, bs, lst, ls = p.get_sort()
    keys = nufft.bin_keys(allom, nf, tile, sizes)
    ref_perm = nufft.stable_sort(keys)
    nkeys = int(np.prod(nufft.tiles_per_dim(nf, tile))) * len(sizes)
    np.testing.assert_array_equal(perm.cpu().numpy(), ref_perm)
    np.testing.assert_array_equal(bs.cpu().numpy(), nufft.bin_start(keys[ref_perm], nkeys))
    lk, lp = nufft.spread_lists(allom, nf, tile, info["w"], sizes)
    np.testing.assert_array_equal(lst.cpu().numpy(), lp)
    np.testing.assert_array_equal(ls.cpu().numpy(), nufft.bin_start(lk, nkeys))


def _ops(dim, C, sense, density, n_trans=None, **kw):
    from hipnufft import MRIHipNUFFT
    from hipnufft.batched import MultiTrajectoryOperator

    trajs = _trajs_2d() if dim == 2 else _trajs_3d()
    if dim == 2:
        trajs = [t for t in trajs if t.shape[0] > 0]  # single operators need samples
    shape = (40, 36) if dim == 2 else (16, 20, 12)
    K = len(trajs)
    smaps = np.stack([gauss_smaps(C, shape, seed=40 + k) for k in range(K)]) if sense else None
    dens = [np.random.default_rng(50 + k).uniform(0.5, 2.0, t.shape[0]).astype(np.float32)
            for k, t in enumerate(trajs)] if density else None
    multi = MultiTrajectoryOperator(trajs, shape, n_coils=C, smaps=smaps, density=dens,
                                    n_trans=n_trans, **kw)
    singles = [MRIHipNUFFT(t, shape, n_coils=C, smaps=None if smaps is None else smaps[k],
                           density=False if dens is None else dens[k], squeeze_dims=False, **kw)
               for k, t in enumerate(trajs)]
    return multi, singles, shape, K


@pytest.mark.parametrize("dim,C,sense,density,n_trans", [
    (2, 3, True, True, None), (2, 3, True, False, 3), (2, 2, False, True, None),
    (3, 4, True, True, None), (3, 4, True, True, 4), (3, 1, False, False, 2)])
def test_multi_matches_single(dim, C, sense, density, n_trans):
    multi, singles, shape, K = _ops(dim, C, sense, density, n_trans)
    ic = () if sense else (C,)
    x = rand_cplx((K, *ic, *shape), seed=7)
    xt = torch.from_numpy(x).to(DEV)
    y = multi.op(xt)
    parts = multi.split(y)
    for k, op in enumerate(singles):
        ref = op.op(xt[k].reshape(1, *(ic or (1,)), *shape)).reshape(C, -1)
        assert rel_l2(parts[k].cpu().numpy(), ref.cpu().numpy()) < 1e-6, k
    yin = torch.from_numpy(rand_cplx(y.shape[0], seed=8)).to(DEV)
    z = multi.adj_op(yin)
    g = multi.data_consistency(xt, yin)
    for k, op in enumerate(singles):
        yk = multi.split(yin)[k].reshape(1, C, -1)
        ref = op.adj_op(yk).reshape(z[k].shape)
        assert rel_l2(z[k].cpu().numpy(), ref.cpu().numpy()) < 1e-6, k
        refg = op.data_consistency(xt[k].reshape(1, *(ic or (1,)), *shape), yk).reshape(g[k].shape)
        assert rel_l2(g[k].cpu().numpy(), refg.cpu().numpy()) < 1e-6, k


def test_multi_chunking_invariant():
    a, _, shape, K = _ops(3, 4, True, True, None)
    b, _, _, _ = _ops(3, 4, True, True, 4)  # one trajectory per chunk
    x = torch.from_numpy(rand_cplx((K, *shape), seed=9)).to(DEV)
    ya, yb = a.op(x), b.op(x)
    assert torch.equal(ya, yb)
    assert torch.equal(a.adj_op(ya), b.adj_op(ya))


def test_multi_cg_matches_single_cg():
    from hipnufft import cg

    multi, singles, shape, K = _ops(2, 3, True, False, None)
    x = torch.from_numpy(rand_cplx((K, *shape), seed=11)).to(DEV)
    y = multi.op(x)
    L = multi.get_lipschitz_cst()
    for graph in (False, True):
        rec = multi.cg(y, max_iter=8, lipschitz=L, graph=graph)
        for k, op in enumerate(singles):
            op.get_lipschitz_cst = (lambda v: (lambda **_: v))(float(L[k]))
            ref = cg(op, multi.split(y)[k].reshape(1, 3, -1), max_iter=8)
            assert rel_l2(rec[k].cpu().numpy(), ref.reshape(rec[k].shape).cpu().numpy()) < 1e-5


def test_multi_lipschitz_matches_single():
    multi, singles, shape, K = _ops(2, 2, True, True, None)
    L = multi.get_lipschitz_cst(max_iter=30, seed=1)
    for k, op in enumerate(singles):
        Ls = op.get_lipschitz_cst(max_iter=30)
        assert abs(L[k] - Ls) / Ls < 2e-2  # independent random starts


def test_dot_rows_bitexact():
    from hipnufft import _capi

    a = torch.from_numpy(rand_cplx((5, 70001), seed=12)).to(DEV)
    b = torch.from_numpy(rand_cplx((5, 70001), seed=13)).to(DEV)
    for conj in (True, False):
        rows = _capi.dot_c64_rows(a, b, conj)
        for r in range(5):
            one = _capi.dot_c64(a[r], b[r], conj)
            assert torch.equal(rows[r], one)
        ref = (np.conj(a.cpu().numpy()) if conj else a.cpu().numpy()).astype(np.complex128)
        ref = np.sum(ref * b.cpu().numpy().astype(np.complex128), axis=1)
        np.testing.assert_allclose(rows[:, 0].cpu().numpy() + 1j * rows[:, 1].cpu().numpy(), ref,
                                   rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("dim,sense", [(2, True), (2, False), (3, True)])
def test_multi_trajectory_vs_oracle(dim, sense):
    """Each problem of the multi-trajectory operator against the exact FP64
    operator of the oracle on its own trajectory (op, adj_op with density,
    data consistency), not only against a second hipnufft operator."""
    from hipnufft.batched import MultiTrajectoryOperator
    from oracle import mri

    trajs = ([t for t in _trajs_2d() if len(t)] if dim == 2 else _trajs_3d())
    shape = (40, 36) if dim == 2 else (16, 20, 12)
    C, K = 3, len(trajs)
    smaps = gauss_smaps(C, shape, 41) if sense else None
    dens = [np.random.default_rng(42 + k).uniform(0.5, 2, len(t)).astype(np.float32)
            for k, t in enumerate(trajs)]
    op = MultiTrajectoryOperator(trajs, shape, n_coils=C, smaps=smaps, density=dens)
    x = rand_cplx(op.img_full_shape, 43)
    ys = [rand_cplx((C, len(t)), 44 + k) for k, t in enumerate(trajs)]
    got_y = op.split(op.op(torch.from_numpy(x).to(DEV)))
    got_x = op.adj_op(ys).cpu().numpy()
    got_dc = op.data_consistency(torch.from_numpy(x).to(DEV), ys).cpu().numpy()
    for k, t in enumerate(trajs):
        # unit samples: the oracle rescales them to radians as the operator does
        # (radians of a trajectory inside [-0.08, 0.08) would be rescaled twice)
        ref = mri.Operator(t, shape, n_coils=C, smaps=smaps, density=dens[k], exact=True)
        xk = x[k]
        assert rel_l2(got_y[k].cpu().numpy(), ref.op(xk.reshape(1, -1, *shape))[0]) < 2e-5, k
        assert rel_l2(got_x[k], ref.adj_op(ys[k].reshape(1, C, -1))[0]) < 2e-5, k
        assert rel_l2(got_dc[k], ref.data_consistency(xk.reshape(1, -1, *shape),
                                                      ys[k].reshape(1, C, -1))[0]) < 2e-5, k
