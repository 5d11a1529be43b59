"""Pins of the ES-kernel NUFFT oracle (CPU).

* type 2 / type 1 against the exact FP64 NDFT at eps=1e-6 (the reference pins
  its NUFFT backends the same way: tests/operators/test_operator_ref.py:49-74);
* exact adjointness of the spread/interp pair and of the full transforms;
* phihat against a brute-force trapezoid integral and, for a Gaussian, the
  closed-form Fourier transform;
* the bin sort against a pure-Python counting sort, the spread lists against
  a brute-force enumeration, and the fold rule on its edge cases;
* kernel parameters: the published FINUFFT table (w, beta) at sigma = 2.
"""

import math

import numpy as np
import pytest

from helpers import radial2d, rand_cplx, rand_traj, rel_l2, truncnorm_traj
from oracle import ndft, nufft


@pytest.mark.parametrize("eps,w,bw", [(1e-2, 3, 2.26), (1e-3, 4, 2.38), (1e-4, 5, 2.30),
                                      (1e-6, 7, 2.30), (1e-7, 8, 2.30)])
def test_kernel_params(eps, w, bw):
    ww, beta = nufft.kernel_params(eps)
    assert ww == w
    assert beta == pytest.approx(bw * w)


def test_next_smooth_even():
    assert [nufft.next_smooth_even(n) for n in (7, 14, 148, 384, 641, 1001)] == [
        8, 16, 150, 384, 648, 1024]


@pytest.mark.parametrize("shape,M,mk", [((256,), 600, "rand"), ((64, 128), 1000, "trunc"),
                                         ((30, 22), 900, "rand"), ((16, 20, 24), 2000, "rand"),
                                         ((64, 64), 5000, "radial")])
def test_nufft_vs_ndft(shape, M, mk):
    d = len(shape)
    k = {"rand": lambda: rand_traj(M, d, 1), "trunc": lambda: truncnorm_traj(M, d, 2),
         "radial": lambda: radial2d(10, 500)}[mk]()
    om = (k * np.float32(2 * np.pi)).astype(np.float32)
    ku = om.astype(np.float64) / (2 * np.pi)
    p = nufft.Plan(om, shape)
    x = rand_cplx(shape, 3)
    c = rand_cplx(len(om), 4)
    y = p.execute(x)
    z = p.execute_adjoint(c)
    assert rel_l2(y, ndft.type2(ku, x, shape)) < 5e-6
    assert rel_l2(z, ndft.type1(ku, c, shape)) < 5e-6
    assert abs(np.vdot(y, c) - np.vdot(x, z)) < 1e-12 * abs(np.vdot(x, z))


def test_isign_plus_is_conjugate():
    shape = (20, 18)
    om = (rand_traj(300, 2, 5) * np.float32(2 * np.pi)).astype(np.float32)
    x = rand_cplx(shape, 6)
    p = nufft.Plan(om, shape, isign=1)
    q = nufft.Plan(om, shape, isign=-1)
    np.testing.assert_allclose(p.execute(x), np.conj(q.execute(np.conj(x))), rtol=1e-12, atol=1e-12)


def test_phihat_brute_and_gaussian():
    w, beta = nufft.kernel_params(1e-6)
    xi = np.linspace(0, np.pi, 5)
    u = np.linspace(-w / 2, w / 2, 400001)
    brute = [np.trapezoid(nufft.es_kernel(u, w, beta) * np.cos(x * u), u) for x in xi]
    np.testing.assert_allclose(nufft.phihat(xi, w, beta), brute, rtol=1e-8)
    # the quadrature itself: Gaussian exp(-u^2/2) on a wide support -> sqrt(2 pi) exp(-xi^2/2)
    xg, wg = np.polynomial.legendre.leggauss(400)
    half = 12.0
    uu = half * (xg + 1) / 2
    val = 2 * np.cos(np.multiply.outer(xi, uu)) @ (np.exp(-uu ** 2 / 2) * wg * half / 2)
    np.testing.assert_allclose(val, np.sqrt(2 * np.pi) * np.exp(-xi ** 2 / 2), rtol=1e-10)


def test_spread_interp_adjoint_pair():
    om = (rand_traj(400, 2, 8) * np.float32(2 * np.pi)).astype(np.float32)
    w, beta = nufft.kernel_params(1e-6)
    nf = (40, 36)
    c = rand_cplx(400, 9).astype(np.complex128)
    g = rand_cplx(nf, 10).astype(np.complex128)
    lhs = np.vdot(nufft.interp(om, g, w, beta), c)
    rhs = np.vdot(g, nufft.spread(om, c, nf, w, beta))
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    # a unit strength spreads the separable kernel: its grid sum is the
    # product of the per-axis kernel sums over the w taps
    one = nufft.spread(om[:1], np.ones(1), nf, w, beta)
    sums = []
    for d in range(2):
        g = nufft.fold(om[0, d], nf[d])
        u = g - (nufft.footprint_start(g, w) + np.arange(w))
        assert np.all(np.abs(u) <= w / 2)
        sums.append(np.sum(nufft.es_kernel(u, w, beta)))
    assert one.sum().real == pytest.approx(sums[0] * sums[1], rel=1e-12)


def _counting_sort(keys, nbins):
    buckets = [[] for _ in range(nbins)]
    for i, k in enumerate(keys):
        buckets[int(k)].append(i)
    return np.array([i for b in buckets for i in b], dtype=np.int64)


@pytest.mark.parametrize("shape,tile", [((64, 64), (32, 32)), ((16, 20, 24), (8, 8, 16)),
                                        ((256,), (512,))])
def test_sort_vs_counting_sort(shape, tile):
    d = len(shape)
    k = np.concatenate([rand_traj(500, d, 11), np.zeros((50, d), np.float32),
                        np.full((20, d), 0.5, np.float32), np.full((20, d), -0.5, np.float32)])
    om = (k * np.float32(2 * np.pi)).astype(np.float32)
    nf = tuple(nufft.fine_size(s, 2.0, 7) for s in shape)
    tile = tuple(min(t, 1 << int(math.log2(n))) for t, n in zip(tile, nf))
    keys = nufft.bin_keys(om, nf, tile)
    nbins = int(np.prod(nufft.tiles_per_dim(nf, tile)))
    assert keys.min() >= 0 and keys.max() < nbins
    np.testing.assert_array_equal(nufft.stable_sort(keys), _counting_sort(keys, nbins))
    bs = nufft.bin_start(keys[nufft.stable_sort(keys)], nbins)
    assert bs[0] == 0 and bs[-1] == len(keys) and np.all(np.diff(bs) >= 0)


def test_spread_lists_brute():
    shape = (16, 12)
    nf = (32, 24)
    tile = (16, 8)
    om = (rand_traj(120, 2, 12) * np.float32(2 * np.pi)).astype(np.float32)
    w = 7
    lk, lp = nufft.spread_lists(om, nf, tile, w)
    perm = nufft.stable_sort(nufft.bin_keys(om, nf, tile))
    nt = nufft.tiles_per_dim(nf, tile)
    brute = []
    for p, o in enumerate(perm):
        cells = []
        for d in range(2):
            g = nufft.fold(om[o, d], nf[d])
            i0 = int(np.ceil(g - w / 2))
            cells.append({((i0 + j) % nf[d]) // tile[d] for j in range(w)})
        for t0 in cells[0]:
            for t1 in cells[1]:
                brute.append((t0 * nt[1] + t1, p))
    brute.sort(key=lambda e: (e[0], e[1]))
    assert list(zip(lk.tolist(), lp.tolist())) == brute


def test_fold_edges():
    nf = 384
    om = np.array([np.pi, -np.pi, -1e-10, 0.0, -0.0, np.float32(np.pi) + 1e-6, 2.9 * np.pi],
                  dtype=np.float32)
    g = nufft.fold(om, nf)
    assert np.all(g >= 0) and np.all(g < nf)
    assert g[3] == 0.0 and g[4] == 0.0
    assert g[0] == pytest.approx(192.0, abs=1e-4) and g[1] == pytest.approx(192.0, abs=1e-4)
    off = nufft.footprint_offset(g, 7)
    assert np.all(off >= 2.5) and np.all(off <= 3.5)


def test_point_sets_partition_sort():
    """Several point sets sorted together: set k occupies sorted positions
    [off_k, off_k+1), and within it the order and bins are exactly those of
    set k sorted alone (shifted by off_k); the lists split the same way."""
    shape, tile, w = (16, 12), (16, 8), 7
    nf = (32, 24)
    sizes = [37, 0, 120, 5]
    rng = np.random.default_rng(5)
    pts = [(rng.uniform(-0.5, 0.5, (m, 2)) * 2 * np.pi).astype(np.float32) for m in sizes]
    om = np.concatenate(pts, 0)
    off = np.concatenate([[0], np.cumsum(sizes)])
    nbins = int(np.prod(nufft.tiles_per_dim(nf, tile)))
    keys = nufft.bin_keys(om, nf, tile, sizes)
    perm = nufft.stable_sort(keys)
    lk, lp = nufft.spread_lists(om, nf, tile, w, sizes)
    for k, m in enumerate(sizes):
        seg = perm[off[k]:off[k + 1]]
        assert np.all((seg >= off[k]) & (seg < off[k + 1]))
        if m == 0:
            continue
        k1 = nufft.bin_keys(pts[k], nf, tile)
        p1 = nufft.stable_sort(k1)
        np.testing.assert_array_equal(seg - off[k], p1)
        np.testing.assert_array_equal(keys[seg] - k * nbins, k1[p1])
        l1k, l1p = nufft.spread_lists(pts[k], nf, tile, w)
        sel = (lk >= k * nbins) & (lk < (k + 1) * nbins)
        np.testing.assert_array_equal(lk[sel] - k * nbins, l1k)
        np.testing.assert_array_equal(lp[sel] - off[k], l1p)
