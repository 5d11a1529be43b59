"""Off-resonance corrected NUFFT in FP64 numpy -- test infrastructure (the oracle).

Restates:

* ``extended_fourier_matrix`` <- ``tests/operators/test_orc.py:20-28``: the
  conjugate-phase model ``F * exp(outer(t, field))`` with the normalised
  Fourier matrix of ``oracle/ndft.py``
* ``op`` / ``adj`` <- ``MRIFourierCorrected.op`` / ``adj_op``
  (``src/mrinufft/operators/off_resonance.py:230-302``): for each
  interpolator l, the NUFFT of ``C_l x`` (smaps per coil) weighted by
  ``B[m % NK, l]``, summed; the adjoint with ``conj(B)``, density and
  ``conj(C_l)`` -- here with the exact NDFT
* ``complex_fieldmap`` <- ``get_complex_fieldmap_rad`` (``extras/field_map.py:176-211``)
* ``histogram`` <- ``_create_histogram`` (``:214-254``), ``numpy.histogramdd``
* ``full_C`` <- ``_full_C`` (``:307-361``)
* ``svd`` / ``mti`` / ``mfi`` <- ``:381-439`` (the full-SVD path), ``:442-496``,
  ``:516-566`` (scikit-learn k-means, as the reference)

Pinned by ``tests/test_oracle_orc.py``: with a field map of L distinct values
and the exact interpolators ``B = exp(t w_l)``, ``C_l = [f == w_l]`` the
operator equals the extended Fourier matrix, forward and adjoint, with
several shots, coils and density; adjointness; and every factorization
reproduces ``exp(t f)`` on the support within the reference's own bound
(``tests/test_offres_exp_approx.py:98-101``: atol = rtol = 5e-3).
"""

from __future__ import annotations

import numpy as np

from . import ndft


def complex_fieldmap(b0_map, r2star_map=None):
    b0 = np.asarray(b0_map)
    if np.iscomplexobj(b0):
        return b0.astype(np.complex128)
    f = 2j * np.pi * b0.astype(np.float64)
    if r2star_map is not None:
        f = f + np.asarray(r2star_map, dtype=np.float64)
    return f


def extended_fourier_matrix(samples, shape, field_map, readout_time):
    """(M, N): ``F[m, n] exp(t_m f_n)`` (normalised F)."""
    F = ndft.fourier_matrix(samples, shape, normalize=True)
    return F * np.exp(np.outer(np.asarray(readout_time, np.float64).ravel(),
                               np.asarray(field_map).ravel()))


def op(samples, shape, x, B, C, smaps=None):
    """x (nb, ic, *shape) -> (nb, nc, M): ``sum_l B[m % NK, l] F(S_c C_l x)[m]``."""
    F = ndft.fourier_matrix(samples, shape, normalize=True)
    M, N = F.shape
    B = np.asarray(B, np.complex128)
    C = np.asarray(C, np.complex128).reshape(B.shape[1], N)
    Bfull = B[np.arange(M) % B.shape[0]]  # (M, L)
    x = np.asarray(x, np.complex128)
    nb = x.shape[0]
    x = x.reshape(nb, -1, N)
    if smaps is not None:
        S = np.asarray(smaps, np.complex128).reshape(-1, N)
        imgs = S[None] * x  # (nb, nc, N)
    else:
        imgs = x
    out = np.zeros((nb, imgs.shape[1], M), np.complex128)
    for l in range(B.shape[1]):
        out += Bfull[None, None, :, l] * ((imgs * C[l]) @ F.T)
    return out


def adj(samples, shape, y, B, C, smaps=None, density=None):
    """y (nb, nc, M) -> (nb, 1 or nc, *shape):
    ``sum_l conj(C_l) sum_c conj(S_c) F^H(d conj(B[., l]) y_c)``."""
    F = ndft.fourier_matrix(samples, shape, normalize=True)
    M, N = F.shape
    B = np.asarray(B, np.complex128)
    C = np.asarray(C, np.complex128).reshape(B.shape[1], N)
    Bfull = B[np.arange(M) % B.shape[0]]
    y = np.asarray(y, np.complex128).reshape(-1, np.shape(y)[-2], M)
    if density is not None:
        y = y * np.asarray(density, np.float64)[None, None]
    out = np.zeros((y.shape[0], y.shape[1], N), np.complex128)
    for l in range(B.shape[1]):
        out += np.conj(C[l]) * ((np.conj(Bfull[:, l])[None, None] * y) @ F.conj())
    if smaps is not None:
        S = np.asarray(smaps, np.complex128).reshape(-1, N)
        out = np.sum(np.conj(S)[None] * out, axis=1, keepdims=True)
    return out.reshape(y.shape[0], out.shape[1], *shape)


def histogram(field_map, mask, n_bins=1000):
    """(counts (nr, ni), complex centres (nr, ni)) of the masked map."""
    f = np.asarray(field_map).astype(np.complex64)
    fm = f[mask] if mask is not None else f.ravel()
    if isinstance(n_bins, (int, np.integer)):
        dR = fm.real.max() - fm.real.min()
        dI = fm.imag.max() - fm.imag.min()
        if dI == 0:
            n_bins = (n_bins, 1)
        elif dR == 0:
            n_bins = (1, n_bins)
        else:
            nr = np.maximum(1, int(np.around(n_bins * dR / dI)))
            ni = np.around(n_bins / nr)
            n_bins = (int(nr), int(ni))
    z = fm.view(np.float32).reshape(-1, 2)
    counts, edges = np.histogramdd(z, bins=n_bins)
    centres = [e[1:] - (e[1] - e[0]) / 2 for e in edges]
    return counts, np.add.outer(centres[0], 1j * centres[1])


def full_C(field_map, c_small, n_bins):
    f = np.asarray(field_map).astype(np.complex64)
    fr, fi = f.real.ravel(), f.imag.ravel()
    idx = []
    for v, n in zip((fr, fi), n_bins):
        lo, hi = v.min(), v.max()
        dv = (hi - lo) / n
        if dv != 0:
            idx.append(np.clip(np.around((v - lo) / dv).astype(int), 0, n - 1))
        else:
            idx.append(np.zeros(v.shape, dtype=int))
    return c_small.reshape(-1, *n_bins)[:, idx[0], idx[1]].reshape(-1, *f.shape)


def _lstsq(a, b):
    """``numpy.linalg.lstsq(a, b, rcond=None)`` as the reference calls it on
    complex64 arrays: singular values below eps32 * max(a.shape) * s_max are
    dropped (computed here in FP64)."""
    return np.linalg.lstsq(a, b, rcond=np.finfo(np.float32).eps * max(a.shape))[0]


def _hist(field_map, mask, n_bins):
    h, w = histogram(field_map, mask, n_bins)
    return h.shape, h.ravel().astype(np.float64), w.ravel().astype(np.complex128)


def svd(field_map, readout_time, mask, L=-1, n_bins=1024):
    hs, h, w = _hist(field_map, mask, n_bins)
    t = np.asarray(readout_time, np.float64).ravel()
    E = np.exp(np.outer(t, w))
    if L == -1:
        L = max(1, int(np.ceil(abs(w[-1] - w[0]) * t.max() / 2 * np.pi)))
    u, _, _ = np.linalg.svd(np.sqrt(h) * E, full_matrices=False)
    B = u[:, :L]
    C = _lstsq(B, E)
    return B, full_C(field_map, C, hs), E


def mti(field_map, readout_time, mask, L=-1, n_bins=1024):
    hs, h, w = _hist(field_map, mask, n_bins)
    t = np.asarray(readout_time, np.float64).ravel()
    if L == -1:
        L = max(1, int(np.ceil(2 * abs(w[-1] - w[0]) * t.max() / np.pi)))
    t_l = np.linspace(t.min(), t.max(), L, dtype=np.float32).astype(np.float64)
    C = np.exp(np.outer(t_l, w))
    E = np.exp(np.outer(t, w))
    B = _lstsq((np.sqrt(h) * C).T, (np.sqrt(h) * E).T)
    return B.T, full_C(field_map, C, hs), E


def mfi(field_map, readout_time, mask, L=9, n_bins=1024):
    from sklearn.cluster import KMeans

    hs, h, w = _hist(field_map, mask, n_bins)
    t = np.asarray(readout_time, np.float64).ravel()
    if L == -1:
        L = max(1, int(np.ceil(4 * abs(w[-1] - w[0])) * t.max() / np.pi))
    pts = w.astype(np.complex64).view(np.float32).reshape(-1, 2)
    w_l = KMeans(n_clusters=L).fit(pts, sample_weight=h).cluster_centers_
    w_l = w_l.astype(np.float32).view(np.complex64).ravel().astype(np.complex128)
    B = np.exp(np.outer(t, w_l))
    E = np.exp(np.outer(t, w))
    C = _lstsq(B, E)
    return B, full_C(field_map, C, hs), E
