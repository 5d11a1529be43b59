"""Stacked 2.5D NUFFT: a 2D NUFFT in (x, y) and a Cartesian FFT along z.

Mirrors ``MRIStackedNUFFTGPU`` (``src/mrinufft/operators/stacked.py:360-838``)
and the helpers ``traj3d2stacked`` / ``stacked2traj3d`` (``:840-900``).  For a
stack-of-stars trajectory (the same 2D pattern on NZ of the Z Cartesian kz
planes) the operator equals the full 3D NUFFT of the equivalent 3D
trajectory, at the cost of a 2D NUFFT per sampled plane.

Device layout: per chunk of T (batch, coil) pairs, the z passes
(``RawStack``, ``libhipnufft``) turn images [.][X][Y][Z] into the sampled
planes [T][NZ][X][Y], which one 2D plan launch transforms as T * NZ
transforms sharing the 2D trajectory; k-space is (B, C, NZ * NS) with the
NS samples of plane i at [i * NS, (i + 1) * NS), as in the reference.
Normalisation follows the reference: ``norm_factor`` is the 2D operator's
times sqrt(2), and the z transform is unitary.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _capi
from ._array import from_device, kind_of, proper_trajectory, to_device
from .base import FourierOperatorBase
from .operator import MRIHipNUFFT


def traj3d2stacked(samples, dim_z, n_samples=0):
    """3D stack trajectory -> (2D trajectory, z index) (``stacked.py:840-868``).

    The z coordinates present (first-appearance order) give the planes; the
    2D pattern is the first ``n_samples`` samples (default: M / planes).
    """
    samples = np.asarray(samples).reshape(-1, 3)
    z_kspace, idx = np.unique(samples[:, 2], return_index=True)
    z_kspace = z_kspace[np.argsort(idx)]
    if n_samples == 0:
        n_samples = int(np.prod(samples.shape[:-1])) // len(z_kspace)
    traj2d = samples[:n_samples, :2]
    z_kspace = np.asarray(proper_trajectory(z_kspace.reshape(-1, 1), "unit")).reshape(-1)
    z_index = np.int32(z_kspace * dim_z + dim_z // 2)
    return traj2d, z_index


def stacked2traj3d(samples2d, z_indexes, dim_z):
    """(2D trajectory, z index) -> flattened 3D trajectory (``stacked.py:871-900``):
    plane i carries kz = (z_index[i] - dim_z // 2) / dim_z."""
    z_kspace = (np.asarray(z_indexes) - dim_z // 2) / dim_z
    locs = np.asarray(proper_trajectory(np.asarray(samples2d), normalize="unit"))
    ns, nz = len(locs), len(z_kspace)
    out = np.zeros((nz, ns, 3), dtype=np.asarray(samples2d).dtype)
    out[:, :, :2] = locs[None]
    out[:, :, 2] = z_kspace[:, None]
    return out.reshape(-1, 3)


def _init_samples(samples, z_index, shape):
    """(2D samples, z positions) from a 3D trajectory with ``z_index="auto"``,
    or a 2D trajectory with a boolean mask / integer list (``stacked.py:135-160``)."""
    samples = samples.detach().cpu().numpy() if isinstance(samples, torch.Tensor) else samples
    samples = np.asarray(samples)
    auto_z = isinstance(z_index, str) and z_index == "auto"
    if samples.shape[-1] == 3 and auto_z:
        return traj3d2stacked(samples, shape[-1])
    if samples.shape[-1] == 2 and not auto_z:
        if z_index is None:
            z_index = np.ones(shape[-1], dtype=bool)
        try:
            zi = np.arange(shape[-1])[z_index]
        except IndexError as e:
            raise ValueError(
                "z-index should be a boolean array of length shape[-1], or an array of integer."
            ) from e
        return samples, np.asarray(zi).reshape(-1)
    raise ValueError("Invalid samples or z-index")


class MRIStackedHipNUFFT(FourierOperatorBase):
    """Stacked NUFFT on MI355X: ``stacked-hipnufft``.

    Parameters follow ``MRIStackedNUFFTGPU`` (``stacked.py:384-400``):
    samples (2D with ``z_index``, or 3D with ``z_index="auto"``); shape
    (X, Y, Z); smaps; n_coils; n_batchs; n_trans ((batch, coil) pairs per
    launch; None: as many as a 4 GiB plane workspace holds); z_index;
    squeeze_dims; smaps_cached (maps are always kept on the device);
    density (for the 2D pattern: False, True / "pipe", array); backend
    ("hipnufft" or an ``MRIHipNUFFT`` 2D operator with squeeze_dims=True, no
    smaps and n_coils = n_trans * len(z_index)); other keyword arguments go
    to the 2D operator.
    """

    backend = "stacked-hipnufft"
    available = MRIHipNUFFT.available
    autograd_available = False

    def __init__(self, samples, shape, smaps=None, n_coils=1, n_batchs=1, n_trans=None,
                 z_index="auto", squeeze_dims=False, smaps_cached=False, density=False,
                 backend="hipnufft", **kwargs):
        super().__init__()
        if len(shape) != 3:
            raise ValueError("stacked operators need a 3D shape (X, Y, Z)")
        self.shape = shape
        self.n_coils = n_coils
        self.n_batchs = n_batchs
        self.squeeze_dims = squeeze_dims
        self.smaps_cached = True
        BC = self.n_batchs * self.n_coils
        if isinstance(backend, str):
            if backend.lower() not in ("hipnufft", "stacked-hipnufft"):
                raise ValueError(f"stacked-hipnufft runs on the hipnufft backend, not {backend}")
            s2d, zi = _init_samples(samples, z_index, shape)
            self._samples2d = np.asarray(s2d).reshape(-1, 2)
            self.z_index = np.asarray(zi, dtype=np.int64)
            NZ = max(len(self.z_index), 1)
            if n_trans is None:
                plane_bytes = 8 * 4 * shape[0] * shape[1] * NZ  # fine 2D grids of one pair
                n_trans = max(1, min(BC, (4 << 30) // max(plane_bytes, 1)))
            self.n_trans = int(n_trans)
            self.operator = MRIHipNUFFT(self._samples2d, tuple(shape[:2]),
                                        n_coils=self.n_trans * NZ, n_trans=self.n_trans * NZ,
                                        smaps=None, squeeze_dims=True, density=density, **kwargs)
        elif isinstance(backend, FourierOperatorBase):
            if tuple(backend.shape) != tuple(shape[:2]):
                raise ValueError("Backend operator should have compatible shape")
            s2d, zi = _init_samples(_unit2d(backend), z_index, shape)
            self._samples2d = np.asarray(s2d).reshape(-1, 2)
            self.z_index = np.asarray(zi, dtype=np.int64)
            NZ = max(len(self.z_index), 1)
            if backend.n_coils % NZ:
                raise ValueError("The backend operator should have ``n_coils * len(z_index)``"
                                 " specified for its coil dimension.")
            if backend.uses_sense:
                raise ValueError("Backend operator should not uses smaps.")
            if not backend.squeeze_dims:
                raise ValueError("Backend operator should have ``squeeze_dims=True``")
            self.operator = backend
            self.n_trans = backend.n_coils // NZ
        else:
            raise ValueError("backend should either be a 2D nufft operator,"
                             " or a str specifying which nufft library to use.")
        self.device = self.operator.device
        self.dtype = np.float32
        self._stack = _capi.RawStack(self.shape, self.z_index)
        self._planes = None
        self._gram = None
        self._smaps = None
        self._smaps_zxy = None
        self.smaps = smaps

    # ------------------------------------------------------------ properties
    @property
    def smaps(self):
        return self._smaps

    @smaps.setter
    def smaps(self, new_smaps):
        """Coil maps (C, X, Y, Z), kept on the device; the coil count follows."""
        self._smaps_zxy = None
        if new_smaps is None:
            self._smaps = None
            return
        s = to_device(new_smaps, torch.complex64, self.device)
        if tuple(s.shape[1:]) != tuple(self.shape):
            raise ValueError("Smaps shape does not match image shape.")
        self.n_coils = s.shape[0]
        self._smaps = s.contiguous()

    @property
    def density(self):
        return self.operator.density

    @density.setter
    def density(self, d):
        self.operator.density = d
        self._gram = None

    @property
    def uses_density(self):
        return self.operator.density is not None

    @property
    def n_samples(self):
        return len(self._samples2d) * len(self.z_index)

    @property
    def norm_factor(self):
        """The 2D operator's factor times sqrt(2) (``stacked.py:447-450``)."""
        return self.operator.norm_factor * math.sqrt(2)

    @property
    def samples(self):
        """(NZ, NS, 3): the 2D samples with kz = z_index / Z - 0.5 (``stacked.py:316-340``)."""
        s2 = np.asarray(self._samples2d, dtype=np.float32)
        out = np.zeros((len(self.z_index), len(s2), 3), dtype=s2.dtype)
        for i, idx in enumerate(self.z_index):
            out[i, :, :2] = s2
            out[i, :, 2] = idx / self.shape[-1] - 0.5
        return out

    @samples.setter
    def samples(self, samples):
        s2d, zi = _init_samples(samples, "auto", self.shape)
        if len(zi) != len(self.z_index):
            raise ValueError("the number of planes cannot change")
        self._samples2d = np.asarray(s2d).reshape(-1, 2)
        self.z_index = np.asarray(zi, dtype=np.int64)
        self._stack = _capi.RawStack(self.shape, self.z_index)
        self.operator.samples = self._samples2d
        self._gram = None

    def _plane_buffer(self, T):
        NZ = len(self.z_index)
        n = T * NZ * self.shape[0] * self.shape[1]
        if self._planes is None or self._planes.numel() < n:
            self._planes = torch.empty(n, dtype=torch.complex64, device=self.device)
        return self._planes[:n].view(T * NZ, self.shape[0], self.shape[1])

    def _chunks(self):
        BC = self.n_batchs * self.n_coils
        for t0 in range(0, BC, self.n_trans):
            yield t0, min(self.n_trans, BC - t0)

    # ------------------------------------------------------------ transforms
    def op(self, data, ksp=None):
        """``y[b, c] = F2D(select(Fz(S_c x_b))) / norm`` -> (B, C, NZ * NS)
        (``stacked.py:527-558``, calibless ``:639-662``)."""
        self.check_shape(image=data, ksp=ksp)
        kind = kind_of(data)
        B, C, NZ, NS = self.n_batchs, self.n_coils, len(self.z_index), len(self._samples2d)
        plan = self.operator.raw_op.plan
        zscale = self.inv_norm_factor / math.sqrt(self.shape[-1])
        with torch.cuda.device(self.device):
            x = to_device(data, torch.complex64, self.device)
            x = x.reshape(B * (1 if self.uses_sense else C), *self.shape)
            out = torch.empty((B * C, NZ, NS), dtype=torch.complex64, device=self.device)
            if NZ:
                for t0, T in self._chunks():
                    planes = self._plane_buffer(T)
                    self._stack.fwd(x, self._smaps, planes, C, t0, T, zscale)
                    plan.sense_op(planes, None, out[t0:t0 + T].view(T * NZ, NS), 1, T * NZ, 1.0)
            out = out.view(B, C, NZ * NS)
        if ksp is not None and isinstance(ksp, torch.Tensor) and ksp.is_cuda:
            ksp.reshape(B, C, NZ * NS).copy_(out)
        return from_device(self._safe_squeeze(out), kind)

    def adj_op(self, coeffs, img=None):
        """``x_b = sum_c conj(S_c) Fz^H(zero-fill(F2D^H(d y[b, c]))) / norm``
        (``stacked.py:711-740``, calibless ``:767-790``)."""
        if img is not None:
            self.check_shape(image=img, ksp=coeffs)
        else:
            self.check_shape(ksp=coeffs)
        kind = kind_of(coeffs)
        B, C, NZ, NS = self.n_batchs, self.n_coils, len(self.z_index), len(self._samples2d)
        plan = self.operator.raw_op.plan
        ic = 1 if self.uses_sense else C
        zscale = 1.0 / math.sqrt(self.shape[-1])
        with torch.cuda.device(self.device):
            y = to_device(coeffs, torch.complex64, self.device).reshape(B * C, NZ, NS)
            out = torch.empty((B, ic, *self.shape), dtype=torch.complex64, device=self.device)
            for t0, T in self._chunks():
                planes = self._plane_buffer(T)
                if NZ:
                    plan.sense_adj(y[t0:t0 + T].reshape(T * NZ, NS), None, self.operator.density,
                                   planes, 1, T * NZ, self.inv_norm_factor)
                self._stack.adj(planes, self._smaps, out, C, t0, T, zscale)
        if img is not None and isinstance(img, torch.Tensor) and img.is_cuda:
            img.reshape(out.shape).copy_(out)
        return from_device(self._safe_squeeze(out), kind)

    # ---------------------------------------------------------------- Gram
    @property
    def toeplitz_exact(self):
        """True when every z plane is sampled exactly once.  Then
        ``A^H D A = norm^-2 S^H (I_z (x) A2^H D A2) S``: the unitary z
        transform commutes with the 2D normal operator acting on each
        z slice, so the Gram form is the 2D Toeplitz operator applied slice
        by slice (no z transform at all)."""
        Z = self.shape[-1]
        return len(self.z_index) == Z and np.array_equal(np.sort(self.z_index), np.arange(Z))

    def compute_toeplitz_kernel(self):
        """The 2D Toeplitz kernel of the stack's normal operator (density
        included); raises when the planes do not cover each z once."""
        if not self.toeplitz_exact:
            raise ValueError("the stacked Toeplitz form needs every z plane sampled exactly once")
        op2 = self.operator
        with torch.cuda.device(self.device):
            self._gram = _capi.RawGram(op2.shape, op2.samples, op2.density, eps=op2.eps,
                                       isign=-1, scale=self.inv_norm_factor ** 2,
                                       n_trans=max(1, self.n_trans * len(self.z_index)))
        return self._gram

    def _maps_zxy(self):
        """The coil maps as (C, Z, X, Y) z-slices, built once per set of maps."""
        if self._smaps_zxy is None:
            self._smaps_zxy = self._smaps.permute(0, 3, 1, 2).contiguous()
        return self._smaps_zxy

    def gram_op(self, data, toeplitz=True):
        """``A^H D A x``: with ``toeplitz`` (and every plane sampled once) the
        2D Toeplitz operator on each z slice of each coil image, expanded by
        and combined with the coil maps; otherwise ``adj_op(op(x))``."""
        self.check_shape(image=data)
        if not (toeplitz and self.toeplitz_exact):
            return self.adj_op(self.op(data))
        kind = kind_of(data)
        B, C = self.n_batchs, self.n_coils
        X, Y, Z = self.shape
        ic = 1 if self.uses_sense else C
        with torch.cuda.device(self.device):
            if self._gram is None:
                self.compute_toeplitz_kernel()
            x = to_device(data, torch.complex64, self.device).reshape(B, ic, X, Y, Z)
            xz = x.permute(0, 1, 4, 2, 3)                             # (B, ic, Z, X, Y)
            v = torch.empty((B, C, Z, X, Y), dtype=torch.complex64, device=self.device)
            if self.uses_sense:
                sz = self._maps_zxy()
                torch.mul(xz, sz[None], out=v)                        # expand by the maps
            else:
                v.copy_(xz)
            g = torch.empty_like(v)
            n2 = B * C * Z
            self._gram.apply(v.view(n2, X, Y), None, g.view(n2, X, Y), n2, 1)
            if self.uses_sense:
                # coil combine by elementwise passes (no reduction kernels,
                # so the form can be captured in a HIP graph)
                acc = torch.mul(g[:, 0:1], torch.conj(sz[0:1])[None])
                for c in range(1, C):
                    acc += g[:, c:c + 1] * torch.conj(sz[c:c + 1])[None]
                g = acc
            out = g.permute(0, 1, 3, 4, 2).contiguous()               # (B, ic, X, Y, Z)
        return from_device(self._safe_squeeze(out), kind)

    def get_lipschitz_cst(self, max_iter=10, **kwargs):
        """Lipschitz constant of the single-coil 2D operator with the density,
        as the reference returns it (``stacked.py:792-836``)."""
        tmp = MRIHipNUFFT(self._samples2d, tuple(self.shape[:2]), n_coils=1, smaps=None,
                          squeeze_dims=True,
                          density=self.operator.density if self.uses_density else False,
                          gpu_device_id=self.device.index or 0)
        return tmp.get_lipschitz_cst(max_iter, **kwargs)


def _unit2d(op):
    """2D samples of an operator in [-0.5, 0.5) (its samples are stored in radians)."""
    s = op.samples
    s = s.detach().cpu().numpy() if isinstance(s, torch.Tensor) else np.asarray(s)
    return s / (2 * np.pi)


class MRIStackedNUFFT(MRIStackedHipNUFFT):
    """``stacked`` (``stacked.py:36-128``): the same operator under the
    generic name and its argument order (samples, shape, backend, smaps,
    z_index, n_coils, n_batchs, squeeze_dims); only the hipnufft backend is
    built."""

    backend = "stacked"

    def __init__(self, samples, shape, backend="hipnufft", smaps=None, z_index="auto", n_coils=1,
                 n_batchs=1, squeeze_dims=False, **kwargs):
        super().__init__(samples, shape, smaps=smaps, n_coils=n_coils, n_batchs=n_batchs,
                         z_index=z_index, squeeze_dims=squeeze_dims, backend=backend, **kwargs)
