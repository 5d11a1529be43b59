"""GPU parity of the raw plan (C ABI) against the CPU oracle.

* bin sort: permutation, bin offsets and spread lists bit-exact against the
  oracle's stable sort (integer work: bit-exact bar);
* type 2 / type 1 (complex64, FP32 accumulation) against the oracle ES-NUFFT
  in FP64 (relative L2 <= 1e-5) and the exact FP64 NDFT (<= 2e-5), eps=1e-6;
* adjointness <Ax, y> = <x, A^H y> to 1e-5;
* SENSE op / adjoint / data consistency against numpy restatements.
"""

import numpy as np
import pytest
import torch

from helpers import gauss_smaps, radial2d, rand_cplx, rand_traj, rel_l2, truncnorm_traj
from oracle import ndft, nufft

pytestmark = pytest.mark.gpu

DEV = "cuda"


def plan_for(traj, shape, n_trans=1, **kw):
    from hipnufft import _capi

    om = torch.from_numpy((traj.astype(np.float32) * np.float32(2 * np.pi)).astype(np.float32))
    p = _capi.RawPlan(shape, n_trans=n_trans, **kw)
    p.setpts(om.to(DEV))
    return p, om.numpy()


CASES = [
    ("1d", (256,), lambda: rand_traj(700, 1, 1)),
    ("2d-asym", (64, 128), lambda: truncnorm_traj(1000, 2, 2)),
    ("2d-radial", (64, 64), lambda: radial2d(10, 500)),
    ("2d-odd", (30, 22), lambda: rand_traj(900, 2, 3)),
    ("3d", (16, 20, 24), lambda: rand_traj(3000, 3, 4)),
    ("3d-radial-center", (32, 32, 32),
     lambda: np.concatenate([rand_traj(1500, 3, 5) * 0.05, rand_traj(500, 3, 6)])),
]


@pytest.mark.parametrize("name,shape,mk", CASES, ids=[c[0] for c in CASES])
def test_sort_bitexact(name, shape, mk):
    traj = mk()
    p, om = plan_for(traj, shape)
    info = p.info()
    D = len(shape)
    nf = info["nf"][3 - D:]
    tile = info["tile"][3 - D:]
    perm, bs, lst, ls = p.get_sort()
    keys = nufft.bin_keys(om, nf, tile)
    ref_perm = nufft.stable_sort(keys)
    nbins = int(np.prod(nufft.tiles_per_dim(nf, tile)))
    np.testing.assert_array_equal(perm.cpu().numpy(), ref_perm)
    np.testing.assert_array_equal(bs.cpu().numpy(), nufft.bin_start(keys[ref_perm], nbins))
    lk, lp = nufft.spread_lists(om, nf, tile, info["w"])
    np.testing.assert_array_equal(lst.cpu().numpy(), lp)
    np.testing.assert_array_equal(ls.cpu().numpy(), nufft.bin_start(lk, nbins))


@pytest.mark.parametrize("split", [0, 64], ids=["whole-tiles", "split-dense-tiles"])
@pytest.mark.parametrize("name,shape,mk", CASES, ids=[c[0] for c in CASES])
def test_type2_type1_vs_oracle(name, shape, mk, split):
    traj = mk()
    p, om = plan_for(traj, shape, spread_max=split, max_subproblem=split)
    x = rand_cplx(shape, 10)
    c = rand_cplx(len(traj), 11)
    y = torch.empty(len(traj), dtype=torch.complex64, device=DEV)
    z = torch.empty(shape, dtype=torch.complex64, device=DEV)
    p.execute(torch.from_numpy(x).to(DEV), y)
    p.execute_adjoint(torch.from_numpy(c).to(DEV), z)
    y = y.cpu().numpy()
    z = z.cpu().numpy()
    ref = nufft.Plan(om, shape)
    assert rel_l2(y, ref.execute(x)) < 1e-5
    assert rel_l2(z, ref.execute_adjoint(c)) < 1e-5
    ku = om.astype(np.float64) / (2 * np.pi)
    assert rel_l2(y, ndft.type2(ku, x, shape)) < 2e-5
    assert rel_l2(z, ndft.type1(ku, c, shape)) < 2e-5
    lhs = np.vdot(y.astype(np.complex128), c)
    rhs = np.vdot(x.astype(np.complex128), z)
    assert abs(lhs - rhs) / abs(rhs) < 1e-5


def test_isign_plus():
    shape = (40, 48)
    traj = rand_traj(800, 2, 7)
    p, om = plan_for(traj, shape, isign=1)
    x = rand_cplx(shape, 12)
    y = torch.empty(len(traj), dtype=torch.complex64, device=DEV)
    p.execute(torch.from_numpy(x).to(DEV), y)
    ku = om.astype(np.float64) / (2 * np.pi)
    ref = np.conj(ndft.type2(ku, np.conj(x), shape))
    assert rel_l2(y.cpu().numpy(), ref) < 2e-5


@pytest.mark.parametrize("n_trans", [1, 3, 8])
def test_sense_op_adj_dc(n_trans):
    shape = (48, 40)
    B, C = 2, 4
    traj = truncnorm_traj(1200, 2, 8)
    p, om = plan_for(traj, shape, n_trans=n_trans, spread_max=100)
    M = len(traj)
    x = rand_cplx((B, *shape), 20)
    s = gauss_smaps(C, shape, 21)
    y = rand_cplx((B, C, M), 22)
    dens = np.random.default_rng(23).uniform(0.5, 2.0, M).astype(np.float32)
    scale = 0.37
    ref = nufft.Plan(om, shape)
    tx, ts, ty, td = (torch.from_numpy(a).to(DEV) for a in (x, s, y, dens))

    out = torch.empty((B, C, M), dtype=torch.complex64, device=DEV)
    p.sense_op(tx, ts, out, B, C, scale)
    exp = np.stack([[scale * ref.execute(x[b] * s[c]) for c in range(C)] for b in range(B)])
    assert rel_l2(out.cpu().numpy(), exp) < 1e-5

    img = torch.empty((B, *shape), dtype=torch.complex64, device=DEV)
    p.sense_adj(ty, ts, td, img, B, C, scale)
    exp = np.stack([scale * sum(np.conj(s[c]) * ref.execute_adjoint(dens * y[b, c])
                                for c in range(C)) for b in range(B)])
    assert rel_l2(img.cpu().numpy(), exp) < 1e-5

    grad = torch.empty((B, *shape), dtype=torch.complex64, device=DEV)
    p.sense_dc(tx, ts, ty, td, grad, B, C, scale)
    exp = np.stack([scale * sum(
        np.conj(s[c]) * ref.execute_adjoint(dens * (scale * ref.execute(x[b] * s[c]) - y[b, c]))
        for c in range(C)) for b in range(B)])
    assert rel_l2(grad.cpu().numpy(), exp) < 1e-5

    # calibrationless: B*C independent transforms
    xc = rand_cplx((B, C, *shape), 24)
    out = torch.empty((B, C, M), dtype=torch.complex64, device=DEV)
    p.sense_op(torch.from_numpy(xc).to(DEV), None, out, B, C, 1.0)
    exp = np.stack([[ref.execute(xc[b, c]) for c in range(C)] for b in range(B)])
    assert rel_l2(out.cpu().numpy(), exp) < 1e-5
    imgc = torch.empty((B, C, *shape), dtype=torch.complex64, device=DEV)
    p.sense_adj(ty, None, None, imgc, B, C, 1.0)
    exp = np.stack([[ref.execute_adjoint(y[b, c]) for c in range(C)] for b in range(B)])
    assert rel_l2(imgc.cpu().numpy(), exp) < 1e-5


def test_dot_fp64():
    from hipnufft import _capi

    a = rand_cplx(100003, 30)
    b = rand_cplx(100003, 31)
    ta, tb = torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV)
    r = _capi.dot_c64(ta, tb, conj=True).cpu().numpy()
    ref = np.vdot(a.astype(np.complex128), b.astype(np.complex128))
    assert abs(complex(r[0], r[1]) - ref) <= 1e-12 * abs(ref) * 100
    r2 = _capi.dot_c64(ta, tb, conj=False).cpu().numpy()
    ref2 = np.dot(a.astype(np.complex128), b.astype(np.complex128))
    assert abs(complex(r2[0], r2[1]) - ref2) <= 1e-10 * abs(ref2)
    # fixed-order reduction: bit-identical on repeat
    r3 = _capi.dot_c64(ta, tb, conj=True).cpu().numpy()
    np.testing.assert_array_equal(r, r3)


@pytest.mark.parametrize("T", [1, 3, 8, 11, 16, 19])
def test_spread_kernels_bit_identical(T):
    """3D 8x8x16 tiles: the cell-owner, block-owner and (>= 8 transforms)
    column-owner spreading kernels sum the same list entries in the same
    order with the same weights: bit-identical grids; the matrix-core kernel
    agrees to FP32 rounding.  Includes split dense tiles (spread_max) and a
    partially filled last transform group."""
    from hipnufft import _capi

    traj = np.concatenate([rand_traj(3000, 3, 41), rand_traj(2500, 3, 42) * 0.06])
    shape = (24, 20, 32)
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    y = torch.from_numpy(rand_cplx((T, traj.shape[0]), seed=43)).to(DEV)
    methods = ["cells", "blocks"] + (["columns"] if T >= 8 else [])
    outs = {}
    for m in methods + ["matrix", "matrix16"]:
        p = _capi.RawPlan(shape, n_trans=T, spread_group=16 if m in ("columns", "matrix16") else 8,
                          spread_max=300,
                          spread_method="matrix" if m.startswith("matrix") else m)
        p.setpts(om)
        assert p.info()["tile"] == (8, 8, 16)
        x = torch.empty((T, *shape), dtype=torch.complex64, device=DEV)
        p.execute_adjoint(y, x)
        outs[m] = x
    torch.cuda.synchronize()
    for m in methods[1:]:
        assert torch.equal(outs["cells"], outs[m]), m
    # matrix cores: same entries, FP32 sums in another order; 16 transforms
    # per workgroup (more than 8 transforms) and 8 give the same bits
    assert rel_l2(outs["matrix"].cpu().numpy(), outs["cells"].cpu().numpy()) < 1e-6
    assert torch.equal(outs["matrix"], outs["matrix16"])
    ref = nufft.Plan(om.cpu().numpy(), shape, eps=1e-6)
    for t in (0, T - 1):
        exp = ref.execute_adjoint(y[t].cpu().numpy().astype(np.complex128))
        assert rel_l2(outs["blocks"][t].cpu().numpy(), exp) < 1e-5


@pytest.mark.parametrize("T", [1, 3, 8, 16, 19, 33])
@pytest.mark.parametrize("tile,split", [(None, 0), (None, 60), ((16, 16), 0), ((8, 64), 90)],
                         ids=["32x32", "32x32-split", "16x16", "8x64-split"])
def test_spread_matrix_2d(tile, split, T):
    """2D matrix-core spreading (blocks of 4 x 16 cells, 8 transforms per
    workgroup) against the cell-owner kernel (<= 1e-6: same list entries and
    weights, FP32 sums in another order) and the FP64 oracle (<= 1e-5),
    with dense tiles split into partial slots, edge tiles past the grid end
    (33 x 50 image), launches of 1 or 3 transforms and a last group of fewer
    than 8 transforms."""
    from hipnufft import _capi

    shape = (33, 50)
    traj = np.concatenate([rand_traj(4000, 2, 61), rand_traj(3000, 2, 62) * 0.05])
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    y = torch.from_numpy(rand_cplx((T, traj.shape[0]), seed=63)).to(DEV)
    kw = dict(n_trans=T, spread_max=split)
    if tile is not None:
        kw["tile"] = tile
    outs = {}
    for m in ("cells", "matrix"):
        p = _capi.RawPlan(shape, spread_method=m, **kw)
        p.setpts(om)
        x = torch.empty((T, *shape), dtype=torch.complex64, device=DEV)
        p.execute_adjoint(y, x)
        outs[m] = x.cpu().numpy()
        p.close()
    assert rel_l2(outs["matrix"], outs["cells"]) < 1e-6
    ref = nufft.Plan(om.cpu().numpy(), shape, eps=1e-6)
    for t in (0, T - 1):
        exp = ref.execute_adjoint(y[t].cpu().numpy().astype(np.complex128))
        assert rel_l2(outs["matrix"][t], exp) < 1e-5


INTERP_KERNELS = [
    ("stream", {}),                                   # every transform of an item per workgroup
    ("stream-small-items", {"max_subproblem": 40}),   # lp = 4 lanes per point, many items
    ("stream-span1", {"interp_span": 1}),
    ("stream-span3", {"interp_span": 3}),
    ("per-transform", {"interp_group": 1}),           # one transform per workgroup
    ("fallback-big-items", {"max_subproblem": 700}),  # items too large for the stream kernel
]


@pytest.mark.parametrize("T", [1, 3, 16, 19])
@pytest.mark.parametrize("shape", [(24, 20, 32), (6, 4, 5)], ids=["grid", "tiny-wrapping"])
def test_interp_kernels_3d(shape, T):
    """3D type 2 through both interpolation kernels (all transforms of a work
    item in one workgroup, staged by direct-to-LDS loads; one transform per
    workgroup) and their work-item splits: each transform against the FP64
    oracle NUFFT (<= 1e-5), and the kernels against each other (<= 1e-6).
    The tiny grid is smaller than the kernel footprint, so regions wrap
    several times around it."""
    from hipnufft import _capi

    traj = np.concatenate([rand_traj(3000, 3, 51), rand_traj(2500, 3, 52) * 0.06])
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    x = torch.from_numpy(rand_cplx((T, *shape), seed=53)).to(DEV)
    outs = {}
    for name, kw in INTERP_KERNELS:
        p = _capi.RawPlan(shape, n_trans=T, **kw)
        p.setpts(om)
        y = torch.empty((T, traj.shape[0]), dtype=torch.complex64, device=DEV)
        p.execute(x, y)
        outs[name] = y.cpu().numpy()
        p.close()
    ref = nufft.Plan(om.cpu().numpy(), shape, eps=1e-6)
    for t in sorted({0, T - 1}):
        exp = ref.execute(x[t].cpu().numpy().astype(np.complex128))
        for name in outs:
            assert rel_l2(outs[name][t], exp) < 1e-5, (name, t)
    for name in outs:
        assert rel_l2(outs[name], outs["per-transform"]) < 1e-6, name


INTERP_KERNELS_2D = [
    ("stream", {}),                                   # every transform of an item per workgroup
    ("stream-small-items", {"interp_group": -1, "max_subproblem": 40}),
    ("stream-span2", {"max_subproblem": 200, "tile": (16, 16), "interp_span": 2}),
    ("grouped", {"interp_group": 4}),                 # 4 transforms per workgroup
    ("fallback-big-items", {"max_subproblem": 700}),  # items too large for the stream kernel
    ("per-transform", {"interp_group": 1}),
]


@pytest.mark.parametrize("T", [1, 3, 16, 19])
@pytest.mark.parametrize("shape", [(72, 40), (6, 5)], ids=["grid", "tiny-wrapping"])
def test_interp_kernels_2d(shape, T):
    """2D type 2 through the streaming kernel (all transforms of a work item
    in one workgroup, the default) and the kernel that stages 1 or 4
    transforms side by side: each transform against the
    FP64 oracle NUFFT (<= 1e-5), and the kernels against each other
    (<= 1e-6).  The tiny grid is smaller than the kernel footprint."""
    from hipnufft import _capi

    traj = np.concatenate([rand_traj(4000, 2, 61), rand_traj(3000, 2, 62) * 0.05])
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    x = torch.from_numpy(rand_cplx((T, *shape), seed=63)).to(DEV)
    outs = {}
    for name, kw in INTERP_KERNELS_2D:
        p = _capi.RawPlan(shape, n_trans=T, **kw)
        p.setpts(om)
        y = torch.empty((T, traj.shape[0]), dtype=torch.complex64, device=DEV)
        p.execute(x, y)
        outs[name] = y.cpu().numpy()
        p.close()
    ref = nufft.Plan(om.cpu().numpy(), shape, eps=1e-6)
    for t in sorted({0, T - 1}):
        exp = ref.execute(x[t].cpu().numpy().astype(np.complex128))
        for name in outs:
            assert rel_l2(outs[name][t], exp) < 1e-5, (name, t)
    for name in outs:
        assert rel_l2(outs[name], outs["per-transform"]) < 1e-6, name


@pytest.mark.parametrize("eps,upsampfac", [(1e-2, 2.0), (1e-4, 2.0), (1e-4, 1.25), (1e-6, 1.5)])
@pytest.mark.parametrize("shape", [(40, 36), (12, 10, 14)], ids=["2d", "3d"])
def test_eps_and_upsampfac_vs_oracle(shape, eps, upsampfac):
    """Kernel widths 3, 5, 7 and 8 (eps 1e-2 .. 1e-6, oversampling 2, 1.25
    and 1.5) through the default kernels: the plan's fine grid and width
    equal the oracle plan's of the same eps and oversampling, and type 2 and
    type 1 of three transforms agree with it to FP32 rounding (<= 1e-5)."""
    from hipnufft import _capi

    d = len(shape)
    T = 3
    traj = np.concatenate([rand_traj(3000, d, 71), rand_traj(1500, d, 72) * 0.1])
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    ref = nufft.Plan(om.cpu().numpy(), shape, eps=eps, upsampfac=upsampfac)
    p = _capi.RawPlan(shape, n_trans=T, eps=eps, upsampfac=upsampfac)
    p.setpts(om)
    info = p.info()
    assert tuple(info["nf"])[-d:] == tuple(ref.nf) and info["w"] == ref.w, (info, ref.nf, ref.w)
    x = rand_cplx((T, *shape), 73)
    c = rand_cplx((T, traj.shape[0]), 74)
    y = torch.empty((T, traj.shape[0]), dtype=torch.complex64, device=DEV)
    z = torch.empty((T, *shape), dtype=torch.complex64, device=DEV)
    p.execute(torch.from_numpy(x).to(DEV), y)
    p.execute_adjoint(torch.from_numpy(c).to(DEV), z)
    p.close()
    y, z = y.cpu().numpy(), z.cpu().numpy()
    for t in (0, T - 1):
        assert rel_l2(y[t], ref.execute(x[t])) < 1e-5, t
        assert rel_l2(z[t], ref.execute_adjoint(c[t])) < 1e-5, t


@pytest.mark.parametrize("T", [5, 12])
@pytest.mark.parametrize("tile,shape,eps", [
    ((8, 8, 8), (16, 12, 10), 1e-6),     # fine grid 32 x 24 x 20: partial edge tiles
    ((4, 8, 16), (20, 16, 24), 1e-6),
    ((4, 4, 16), (12, 14, 18), 1e-3),    # w = 4, a last-axis tile past the grid end
    ((8, 4, 8), (6, 6, 6), 1e-2),        # w = 3, nf = 12: one block per wave, wraps often
])
def test_spread_blocks_other_tiles(tile, shape, eps, T):
    """Block-owner spreading on other tile shapes (1 or 2 blocks per wave,
    tiles cut by the grid end, small grids where footprints wrap): bit-equal
    to the cell-owner kernel and within eps of the FP64 oracle.  With 12
    transforms the matrix-core kernel takes 16 per workgroup (waves without
    a block idle)."""
    from hipnufft import _capi

    traj = np.concatenate([rand_traj(1500, 3, 44), rand_traj(800, 3, 45) * 0.1])
    om = torch.from_numpy((traj * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    y = torch.from_numpy(rand_cplx((T, traj.shape[0]), seed=46)).to(DEV)
    outs = {}
    for m in ("cells", "blocks", "matrix"):
        p = _capi.RawPlan(shape, n_trans=T, eps=eps, tile=tile, spread_max=200, spread_method=m)
        p.setpts(om)
        assert p.info()["tile"] == tile
        x = torch.empty((T, *shape), dtype=torch.complex64, device=DEV)
        p.execute_adjoint(y, x)
        outs[m] = x
    torch.cuda.synchronize()
    assert torch.equal(outs["cells"], outs["blocks"])
    assert rel_l2(outs["matrix"].cpu().numpy(), outs["cells"].cpu().numpy()) < 1e-6
    ref = nufft.Plan(om.cpu().numpy(), shape, eps=eps)
    exp = ref.execute_adjoint(y[T - 1].cpu().numpy().astype(np.complex128))
    assert rel_l2(outs["blocks"][T - 1].cpu().numpy(), exp) < 10 * eps


def test_spread_method_rejected_where_unsupported():
    from hipnufft import _capi

    om = torch.from_numpy((rand_traj(500, 2, 47) * np.float32(2 * np.pi)).astype(np.float32)).to(DEV)
    y = torch.from_numpy(rand_cplx((1, 500), seed=48)).to(DEV)
    p = _capi.RawPlan((32, 32), n_trans=1, spread_method="blocks")
    p.setpts(om)
    x = torch.empty((1, 32, 32), dtype=torch.complex64, device=DEV)
    with pytest.raises(ValueError, match="block spreading"):
        p.execute_adjoint(y, x)
    with pytest.raises(ValueError):
        _capi.RawPlan((32, 32), n_trans=1, spread_method="fastest")
