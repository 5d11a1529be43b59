"""Exponential-of-semicircle NUFFT in numpy -- test infrastructure (the oracle).

The reference delegates all NUFFT arithmetic to finufft / cufinufft >= 2.5.0
(``pyproject.toml:21,23``), which are not vendored under the reference tree
and are not importable here.  Their call sites are
``operators/interfaces/finufft.py:43-76`` (type 2 plan, ``execute`` /
``execute_adjoint``) and ``operators/interfaces/cufinufft.py:83-133``.  This
module restates the published algorithm of those libraries (Barnett, Magland
& af Klinteberg, "A parallel non-uniform fast Fourier transform library based
on an exponential of semicircle kernel", SIAM J. Sci. Comput. 41(5), 2019;
kernel definition also in the reference's ``docs/explanations/nufft.rst:255-307``):

* kernel ``phi(u) = exp(beta (sqrt(1 - (2u/w)^2) - 1))`` on ``|u| <= w/2``
  (``u`` in fine-grid cells);
* width ``w = ceil(-log10(eps/10))`` and ``beta = 2.30 w`` for upsampling
  factor 2 (2.20 / 2.26 / 2.38 for w = 2 / 3 / 4), otherwise
  ``w = ceil(-ln(eps) / (pi sqrt(1 - 1/sigma)))`` and
  ``beta = 0.97 pi (1 - 1/(2 sigma)) w``;
* fine grid ``nf = next 2^a 3^b 5^c even >= max(int(sigma N), 2 w)``;
* type 2: deconvolve by ``1/phihat(2 pi k / nf)``, zero-pad, FFT, interpolate;
  type 1 is the exact adjoint: spread, FFT, deconvolve, crop.

The fold / bin / footprint rules (``fold``, ``bin_keys``, ``footprint_start``)
are this project's own definitions, evaluated in IEEE double; the HIP kernels
use the very same double operations, so the bin sort is checked bit-exact.

Pinned by ``tests/test_oracle_nufft.py``: type 2 / type 1 against the exact
FP64 NDFT of ``oracle.ndft`` (relative L2 ~1e-6 at eps = 1e-6, the accuracy
the FINUFFT paper states); exact adjointness; ``phihat`` against a
closed-form Fourier transform check on a Gaussian-like kernel and a brute
trapezoid sum; the stable sort against ``numpy.argsort(kind="stable")`` and a
pure-Python counting sort; the spread lists against brute-force enumeration;
several point sets sorted together against each set sorted alone.
"""

from __future__ import annotations

import math

import numpy as np


# ----------------------------------------------------------------------------
# kernel parameters and grid sizes
# ----------------------------------------------------------------------------
def kernel_params(eps: float, upsampfac: float = 2.0) -> tuple[int, float]:
    """Return (w, beta) for tolerance ``eps`` (published FINUFFT choice)."""
    if upsampfac == 2.0:
        w = math.ceil(-math.log10(eps / 10.0))
    else:
        w = math.ceil(-math.log(eps) / (math.pi * math.sqrt(1.0 - 1.0 / upsampfac)))
    w = int(min(max(w, 2), 16))
    if upsampfac == 2.0:
        beta_over_w = {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30)
    else:
        beta_over_w = 0.97 * math.pi * (1.0 - 1.0 / (2.0 * upsampfac))
    return w, beta_over_w * w


def next_smooth_even(n: int) -> int:
    """Smallest even integer >= n whose only prime factors are 2, 3, 5."""
    n = max(int(n), 2)
    if n % 2:
        n += 1
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 2


def fine_size(N: int, upsampfac: float, w: int) -> int:
    """Oversampled grid size along one axis."""
    nf = int(upsampfac * N)
    nf = max(nf, 2 * w)
    return next_smooth_even(nf)


def es_kernel(u, w, beta):
    """ES kernel value at offset ``u`` (fine-grid cells)."""
    z = 2.0 * np.asarray(u, dtype=np.float64) / w
    arg = 1.0 - z * z
    out = np.exp(beta * (np.sqrt(np.maximum(arg, 0.0)) - 1.0))
    return np.where(arg >= 0.0, out, 0.0)


def phihat(xi, w, beta, nquad=400):
    """Fourier transform ``int phi(u) exp(i xi u) du`` by Gauss-Legendre."""
    x, wt = np.polynomial.legendre.leggauss(nquad)
    half = w / 2.0
    u = half * (x + 1.0) / 2.0  # nodes on [0, w/2]
    wt = wt * half / 2.0
    xi = np.asarray(xi, dtype=np.float64)
    return 2.0 * np.cos(np.multiply.outer(xi, u)) @ (es_kernel(u, w, beta) * wt)


def mode_index(N):
    """Integer frequency of image index i: ``k = i - N // 2``."""
    return np.arange(N) - N // 2


def deconv_factors(N, nf, w, beta):
    """Per-axis deconvolution ``1/phihat(2 pi k / nf)`` in image-index order."""
    return 1.0 / phihat(2.0 * np.pi * mode_index(N) / nf, w, beta)


def grid_positions(N, nf):
    """Fine-grid position of image index i: ``k mod nf``."""
    return mode_index(N) % nf


# ----------------------------------------------------------------------------
# FP64 fold / bins / footprint (bit-exact with the GPU)
# ----------------------------------------------------------------------------
def fold_scale(nf):
    """Factor radians -> fine-grid cells, in double."""
    return nf / (2.0 * math.pi)


def fold(omega, nf):
    """Fold float32 radians into a double grid coordinate in [0, nf).

    One IEEE double multiply and at most two conditional adds per branch, so
    the result is reproducible bit for bit on any IEEE machine.
    """
    nff = float(nf)
    g = np.asarray(omega, dtype=np.float32).astype(np.float64) * fold_scale(nf)
    for _ in range(2):
        g = np.where(g < 0, g + nff, g)
    for _ in range(2):
        g = np.where(g >= nff, g - nff, g)
    return g


def footprint_start(g, w):
    """First footprint cell ``ceil(g - w/2)`` (double arithmetic)."""
    return np.ceil(np.asarray(g, dtype=np.float64) - w / 2.0).astype(np.int64)


def footprint_offset(g, w):
    """float32 offset ``g - i0`` in [w/2 - 1, w/2] used by the GPU taps."""
    g = np.asarray(g, dtype=np.float64)
    return (g - footprint_start(g, w)).astype(np.float32)


def tiles_per_dim(nf, tile):
    return [(n + t - 1) // t for n, t in zip(nf, tile)]


def set_ids(M, set_sizes=None):
    """Point-set index of each of the M points (sets stored back to back)."""
    if set_sizes is None:
        return np.zeros(M, dtype=np.int64)
    ids = np.repeat(np.arange(len(set_sizes), dtype=np.int64), np.asarray(set_sizes))
    assert ids.shape[0] == M
    return ids


def bin_keys(omega, nf, tile, set_sizes=None):
    """C-order tile key of the tile containing each point; with several point
    sets (``set_sizes``, sets back to back) the key is set * ntiles + tile."""
    omega = np.asarray(omega, dtype=np.float32)
    nt = tiles_per_dim(nf, tile)
    key = np.zeros(omega.shape[0], dtype=np.int64)
    for d in range(len(nf)):
        g = fold(omega[:, d], nf[d])
        b = np.floor(g * (1.0 / tile[d])).astype(np.int64)
        key = key * nt[d] + b
    return key + set_ids(omega.shape[0], set_sizes) * int(np.prod(nt))


def stable_sort(keys):
    """Stable counting-sort permutation (sorted position -> original index)."""
    return np.argsort(np.asarray(keys), kind="stable")


def bin_start(sorted_keys, nbins):
    """Offsets of each bin's first point in sorted order (length nbins + 1)."""
    return np.searchsorted(sorted_keys, np.arange(nbins + 1), side="left")


def footprint_tiles(g, nf, tile, w):
    """Distinct tiles touched by one point's footprint along one axis."""
    i0 = int(footprint_start(g, w))
    out = []
    for j in range(w):
        t = ((i0 + j) % nf) // tile
        if t not in out:
            out.append(t)
    return out


def spread_lists(omega, nf, tile, w, set_sizes=None):
    """Tile -> sorted-point lists with duplicates (for gather spreading).

    Returns (list_keys, list_points): for every (point, touched tile) pair
    generated in sorted-point order, the tile key (set * ntiles + tile with
    several point sets) and the sorted position of the point, then stably
    sorted by key.
    """
    omega = np.asarray(omega, dtype=np.float32)
    keys = bin_keys(omega, nf, tile, set_sizes)
    perm = stable_sort(keys)
    nt = tiles_per_dim(nf, tile)
    sid = set_ids(omega.shape[0], set_sizes)[perm]
    ntiles = int(np.prod(nt))
    g = np.stack([fold(omega[:, d], nf[d]) for d in range(len(nf))], 1)[perm]
    lk, lp = [], []
    for p in range(len(perm)):
        per = [footprint_tiles(g[p, d], nf[d], tile[d], w) for d in range(len(nf))]
        combos = [0]
        for d in range(len(nf)):
            combos = [c * nt[d] + t for c in combos for t in per[d]]
        for c in combos:
            lk.append(int(sid[p]) * ntiles + c)
            lp.append(p)
    lk = np.asarray(lk, dtype=np.int64)
    lp = np.asarray(lp, dtype=np.int64)
    o = stable_sort(lk)
    return lk[o], lp[o]


# ----------------------------------------------------------------------------
# spreading / interpolation / transforms (float64 arithmetic)
# ----------------------------------------------------------------------------
def _axis_weights(g, nfd, w, beta):
    i0 = footprint_start(g, w)
    j = np.arange(w)
    cells = (i0[:, None] + j[None, :]) % nfd
    u = g[:, None] - (i0[:, None] + j[None, :])
    return cells, es_kernel(u, w, beta)


def _flat_taps(omega, nf, w, beta):
    omega = np.asarray(omega, dtype=np.float32)
    flat = np.zeros((omega.shape[0], 1), dtype=np.int64)
    wt = np.ones((omega.shape[0], 1))
    for d in range(len(nf)):
        cells, wd = _axis_weights(fold(omega[:, d], nf[d]), nf[d], w, beta)
        flat = (flat[:, :, None] * nf[d] + cells[:, None, :]).reshape(len(omega), -1)
        wt = (wt[:, :, None] * wd[:, None, :]).reshape(len(omega), -1)
    return flat, wt


def spread(omega, c, nf, w, beta, chunk=20000):
    """Spread strengths ``c`` onto a periodic grid of shape ``nf``."""
    c = np.asarray(c, dtype=np.complex128).reshape(-1)
    G = int(np.prod(nf))
    re = np.zeros(G)
    im = np.zeros(G)
    for s in range(0, len(c), chunk):
        flat, wt = _flat_taps(omega[s : s + chunk], nf, w, beta)
        v = wt * c[s : s + chunk, None]
        re += np.bincount(flat.ravel(), weights=v.real.ravel(), minlength=G)
        im += np.bincount(flat.ravel(), weights=v.imag.ravel(), minlength=G)
    return (re + 1j * im).reshape(nf)


def interp(omega, grid, w, beta, chunk=20000):
    """Interpolate a periodic grid at the points."""
    nf = grid.shape
    gf = grid.reshape(-1)
    out = np.empty(len(omega), dtype=np.complex128)
    for s in range(0, len(omega), chunk):
        flat, wt = _flat_taps(omega[s : s + chunk], nf, w, beta)
        out[s : s + chunk] = np.sum(wt * gf[flat], axis=1)
    return out


class Plan:
    """Type-2 style plan executing both directions, like finufft's ``Plan``.

    ``execute`` is type 2 (sign ``isign``, default -1) and
    ``execute_adjoint`` its adjoint (type 1, sign ``-isign``), mirroring
    ``RawFinufftPlan.op`` / ``adj_op`` (``interfaces/finufft.py:64-76``).
    ``spreadinterponly`` mirrors the ``spreadinterponly=1`` option used by
    Pipe (``interfaces/finufft.py:225-232``): the grid is the image itself
    (size ``shape``), no FFT and no deconvolution.
    """

    def __init__(self, omega, shape, eps=1e-6, upsampfac=2.0, isign=-1,
                 spreadinterponly=False):
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.omega = np.asarray(omega, dtype=np.float32).reshape(-1, self.ndim)
        self.w, self.beta = kernel_params(eps, upsampfac)
        self.isign = isign
        self.sio = spreadinterponly
        if spreadinterponly:
            self.nf = self.shape
        else:
            self.nf = tuple(fine_size(n, upsampfac, self.w) for n in self.shape)
            self.p = [deconv_factors(n, f, self.w, self.beta)
                      for n, f in zip(self.shape, self.nf)]
            self.pos = [grid_positions(n, f) for n, f in zip(self.shape, self.nf)]

    def _deconv(self):
        out = np.ones(self.shape)
        for d, p in enumerate(self.p):
            sh = [1] * self.ndim
            sh[d] = -1
            out = out * p.reshape(sh)
        return out

    def execute(self, image):
        x = np.asarray(image, dtype=np.complex128).reshape(self.shape)
        if self.sio:
            return interp(self.omega, x, self.w, self.beta)
        grid = np.zeros(self.nf, dtype=np.complex128)
        grid[np.ix_(*self.pos)] = x * self._deconv()
        if self.isign < 0:
            grid = np.fft.fftn(grid)
        else:
            grid = np.fft.ifftn(grid) * grid.size
        return interp(self.omega, grid, self.w, self.beta)

    def execute_adjoint(self, coeffs):
        grid = spread(self.omega, coeffs, self.nf, self.w, self.beta)
        if self.sio:
            return grid
        if self.isign < 0:
            grid = np.fft.ifftn(grid) * grid.size
        else:
            grid = np.fft.fftn(grid)
        return grid[np.ix_(*self.pos)] * self._deconv()
