"""Field maps and off-resonance interpolators (``src/mrinufft/extras/field_map.py``).

The off-resonance term ``exp(t_m * f_n)`` (``f = R2* + 2j pi B0`` in rad/s)
is approximated by ``sum_l B[m, l] C[l, n]`` with L temporal (B) and spatial
(C) interpolators, computed from a histogram of the field map:

* ``svd`` (``:381-439``): B = the L leading left singular vectors of the
  histogram-weighted ``E = exp(t w_k)``, C = least squares of B against E;
* ``mti`` (``:442-496``): C = ``exp(t_l w_k)`` at L equispaced times, B by
  weighted least squares;
* ``mfi`` (``:516-566``): B = ``exp(t w_l)`` at L k-means centres of the
  weighted histogram (scikit-learn, on the host as in the reference), C by
  least squares.

The histogram, the exponentials, the SVD and the least squares run with
torch in complex128 on the field map's device (the CUDA device when the
input is a host array and a GPU is present); results are complex64 in the
caller's array kind.  C maps every voxel to its histogram bin like
``_full_C`` (``:307-361``).  ``make_b0map`` / ``make_t2smap`` restate the
synthetic maps of ``:30-120``.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from ._array import from_device, kind_of
from ._registry import MethodRegister

register_orc = MethodRegister("orc_factorization")
get_orc_factorization = register_orc.make_getter()

_EPS32 = float(np.finfo(np.float32).eps)


# ---------------------------------------------------------------- dummy maps
def _radius(shape):
    axes = [(np.arange(n) - n // 2) / n for n in shape]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.sqrt(sum(g ** 2 for g in grids))


def make_b0map(shape, b0range=(-300, 300), mask=None):
    """Radial B0 map in Hz within ``b0range`` and its support (``field_map.py:30-68``).

    The map grows with the normalised radius from the image centre; the
    support is the disk / sphere of radius 0.3 unless ``mask`` is given.
    """
    rr = _radius(shape)
    if mask is None:
        mask = rr < 0.3
    fieldmap = rr * mask
    fieldmap = (b0range[1] - b0range[0]) * fieldmap / fieldmap.max() + b0range[0]
    fieldmap *= mask
    fieldmap = np.nan_to_num(fieldmap, neginf=0.0, posinf=0.0)
    return fieldmap.astype(np.float32), mask


def make_t2smap(shape, t2svalue=15.0, mask=None):
    """Homogeneous T2* map in ms on the support (``field_map.py:71-107``)."""
    if mask is None:
        mask = _radius(shape) < 0.3
    fieldmap = np.nan_to_num(t2svalue * mask, neginf=0.0, posinf=0.0)
    return fieldmap.astype(np.float32), mask


# ---------------------------------------------------------------- helpers
def _device_for(x):
    if isinstance(x, torch.Tensor):
        return x.device
    return torch.device("cuda", 0) if torch.cuda.is_available() else torch.device("cpu")


def _as_tensor(x, device, dtype=None):
    if x is None:
        return None
    t = x.detach() if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
    t = t.to(device)
    return t.to(dtype) if dtype is not None else t


def get_complex_fieldmap_rad(b0_map, r2star_map=None):
    """``R2* + 2j pi B0`` in rad/s, complex64 (``field_map.py:176-211``).

    A complex ``b0_map`` is returned unchanged.
    """
    if isinstance(b0_map, torch.Tensor):
        if b0_map.is_complex():
            return b0_map
        f = (2j * math.pi) * b0_map.to(torch.float32).to(torch.complex64)
        if r2star_map is not None:
            f = f + torch.as_tensor(r2star_map, dtype=torch.float32, device=b0_map.device)
        return f.to(torch.complex64)
    b0 = np.asarray(b0_map)
    if np.iscomplexobj(b0):
        return b0
    f = np.complex64(2j * np.pi) * b0.astype(np.float32, copy=False)
    if r2star_map is not None:
        f = f + np.asarray(r2star_map, dtype=np.float32)
    return f.astype(np.complex64, copy=False)


def _bin_counts(n_bins, d_re, d_im):
    if not isinstance(n_bins, (int, np.integer)):
        return tuple(int(b) for b in n_bins)
    if d_im == 0:
        return (int(n_bins), 1)
    if d_re == 0:
        return (1, int(n_bins))
    nr = max(1, int(np.around(np.float32(n_bins) * np.float32(d_re) / np.float32(d_im))))
    ni = int(np.around(n_bins / nr))
    return (nr, ni)


def _edges(v, nb):
    lo, hi = float(v.min()), float(v.max())
    if lo == hi:
        lo, hi = lo - 0.5, hi + 0.5
    return torch.linspace(lo, hi, nb + 1, dtype=torch.float64, device=v.device)


def _create_histogram(field_map, mask, n_bins=1000):
    """(counts (nr, ni), complex bin centres (nr, ni)) of the masked field map
    (``field_map.py:214-254``; ``numpy.histogramdd`` bins: equal widths over
    [min, max], last bin closed)."""
    f = field_map[mask] if mask is not None else field_map.reshape(-1)
    re = f.real.to(torch.float64)
    im = f.imag.to(torch.float64)
    nb = _bin_counts(n_bins, float(f.real.max() - f.real.min()),
                     float(f.imag.max() - f.imag.min()))
    idx, centres = [], []
    for v, n in zip((re, im), nb):
        e = _edges(v, n)
        i = torch.searchsorted(e, v, right=True) - 1
        idx.append(i.clamp(0, n - 1))  # the maximum lands in the last bin
        centres.append(e[1:] - (e[1] - e[0]) / 2)
    counts = torch.bincount(idx[0] * nb[1] + idx[1], minlength=nb[0] * nb[1])
    counts = counts.to(torch.float64).reshape(nb)
    w = centres[0][:, None] + 1j * centres[1][None, :]
    return counts, w.to(torch.complex128)


class _CLazy:
    """Spatial interpolators kept as (L, nr, ni) bin values plus each voxel's
    bin (``field_map.py:257-304``); ``C[l]`` expands interpolator l."""

    def __init__(self, c_small, idxr, idxi, img_shape, kind):
        self.C_small = c_small
        self.idxr = idxr
        self.idxi = idxi
        self.img_shape = tuple(img_shape)
        self._kind = kind

    def __len__(self):
        return len(self.C_small)

    @property
    def shape(self):
        return (len(self), *self.img_shape)

    def full(self):
        """All interpolators as one (L, *shape) complex64 device tensor."""
        return self.C_small[:, self.idxr, self.idxi].reshape(-1, *self.img_shape)

    def __getitem__(self, i):
        return from_device(self.C_small[i][self.idxr, self.idxi].reshape(self.img_shape),
                           self._kind)


def _full_C(field_map, c_small, n_bins, lazy, kind):
    """Each voxel takes the interpolator values of its bin: index
    ``round((f - min) / ((max - min) / nbins))`` over the whole map, clipped
    (``field_map.py:307-361``)."""
    # float32 arithmetic, as the reference applies it to the complex64 map
    fr = field_map.real.reshape(-1).to(torch.float32)
    fi = field_map.imag.reshape(-1).to(torch.float32)
    idx = []
    for v, n in zip((fr, fi), n_bins):
        lo, hi = v.min(), v.max()
        dv = (hi - lo) / n
        if float(dv) != 0.0:
            idx.append(torch.round((v - lo) / dv).to(torch.int64).clamp(0, n - 1))
        else:
            idx.append(torch.zeros_like(v, dtype=torch.int64))
    c_sr = c_small.to(torch.complex64).reshape(-1, *n_bins)
    lazy_c = _CLazy(c_sr, idx[0], idx[1], field_map.shape, kind)
    return lazy_c if lazy else lazy_c.full()


def _lstsq(a, b):
    """Minimum-norm least squares through the SVD, singular values below
    ``eps32 * max(a.shape) * s_max`` dropped (``numpy.linalg.lstsq`` with
    ``rcond=None`` on complex64 input)."""
    u, s, vh = torch.linalg.svd(a, full_matrices=False)
    keep = s > _EPS32 * max(a.shape) * s[0]
    sinv = torch.where(keep, 1.0 / torch.where(keep, s, torch.ones_like(s)), torch.zeros_like(s))
    return vh.conj().T @ (sinv[:, None].to(a.dtype) * (u.conj().T @ b))


def _prepare(field_map, readout_time, mask):
    kind = kind_of(field_map)
    dev = _device_for(field_map)
    f = _as_tensor(get_complex_fieldmap_rad(field_map), dev).to(torch.complex64)
    t = _as_tensor(readout_time, dev, torch.float32).reshape(-1)
    m = _as_tensor(mask, dev, torch.bool) if mask is not None else None
    return kind, f, t, m


def _out(kind, B, C, E):
    B = from_device(B.to(torch.complex64), kind)
    C = C if isinstance(C, _CLazy) else from_device(C, kind)
    return B, C, from_device(E.to(torch.complex64), kind)


@register_orc("svd")
def compute_svd_coefficients(field_map, readout_time, mask, L=-1, n_bins=1024, lazy=False,
                             partial_svd=True):
    """Truncated-SVD interpolators (``field_map.py:381-439``): B (Nt, L), C (L, *shape), E.

    ``partial_svd`` is accepted for compatibility; the SVD is always the full
    thin SVD of the (Nt, nbins) matrix, truncated to L.
    """
    kind, f, t, m = _prepare(field_map, readout_time, mask)
    h, w = _create_histogram(f, m, n_bins)
    hist_shape = tuple(h.shape)
    h, w = h.reshape(-1), w.reshape(-1)
    t64 = t.to(torch.float64)
    E = torch.exp(t64[:, None].to(torch.complex128) * w[None, :])
    Ew = torch.sqrt(h)[None, :].to(torch.complex128) * E
    if L == -1:
        L = max(1, int(math.ceil(abs(complex(w[-1] - w[0])) * float(t.max()) / 2 * math.pi)))
    u, _, _ = torch.linalg.svd(Ew, full_matrices=False)
    B = u[:, :L]
    C = _lstsq(B, E)
    return _out(kind, B, _full_C(f, C, hist_shape, lazy, kind), E)


@register_orc("mti")
def compute_mti_coefficients(field_map, readout_time, mask, L=-1, n_bins=1024, lazy=False):
    """Mixed-time interpolators (``field_map.py:442-496``)."""
    kind, f, t, m = _prepare(field_map, readout_time, mask)
    h, w = _create_histogram(f, m, n_bins)
    hist_shape = tuple(h.shape)
    h, w = h.reshape(-1), w.reshape(-1)
    if L == -1:
        L = max(1, int(math.ceil(2 * abs(complex(w[-1] - w[0])) * float(t.max()) / math.pi)))
    t_l = torch.linspace(float(t.min()), float(t.max()), L, dtype=torch.float32,
                         device=t.device).to(torch.float64)
    t64 = t.to(torch.float64)
    C = torch.exp(t_l[:, None].to(torch.complex128) * w[None, :])
    E = torch.exp(t64[:, None].to(torch.complex128) * w[None, :])
    sh = torch.sqrt(h)[None, :].to(torch.complex128)
    B = _lstsq((sh * C).T, (sh * E).T)
    return _out(kind, B.T, _full_C(f, C, hist_shape, lazy, kind), E)


def _cluster_centres(w, h, L):
    """L k-means centres of the bin centres weighted by the counts (host,
    scikit-learn; ``field_map.py:499-513``)."""
    try:
        from sklearn.cluster import KMeans
    except ImportError as err:
        raise ImportError(
            "The scikit-learn module is not available. Please install "
            "it along with the [extra] dependencies "
            "or using `pip install scikit-learn`.") from err
    pts = w.to(torch.complex64).cpu().numpy().view(np.float32).reshape(-1, 2)
    km = KMeans(n_clusters=L).fit(pts, sample_weight=h.cpu().numpy())
    return km.cluster_centers_.astype(np.float32).view(np.complex64).reshape(-1)


@register_orc("mfi")
def compute_mfi_coefficients(field_map, readout_time, mask, L=9, n_bins=1024, lazy=True):
    """Multi-frequency interpolators (``field_map.py:516-566``)."""
    kind, f, t, m = _prepare(field_map, readout_time, mask)
    h, w = _create_histogram(f, m, n_bins)
    hist_shape = tuple(h.shape)
    h, w = h.reshape(-1), w.reshape(-1)
    if L == -1:
        L = max(1, int(math.ceil(4 * abs(complex(w[-1] - w[0]))) * float(t.max()) / math.pi))
    w_l = torch.from_numpy(_cluster_centres(w, h, L)).to(t.device).to(torch.complex128)
    t64 = t.to(torch.float64)[:, None].to(torch.complex128)
    B = torch.exp(t64 * w_l[None, :])
    E = torch.exp(t64 * w[None, :])
    C = _lstsq(B, E)
    return _out(kind, B, _full_C(f, C, hist_shape, lazy, kind), E)
