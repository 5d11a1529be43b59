"""The ``hipnufft`` MRI operator: a drop-in for mri-nufft's GPU backends.

Mirrors the constructor and surface of ``MRICufiNUFFT``
(``src/mrinufft/operators/interfaces/cufinufft.py:140-1331``): SENSE and
calibrationless operators, density compensation, batching, normalisation,
``data_consistency``, ``get_lipschitz_cst``, ``pinv_solver``, ``pipe``,
``update_samples`` and the autograd hooks.  All arithmetic runs in
``libhipnufft.so``; this module only moves device pointers.

Differences by design (no behavioural change for callers):

* all coils of all batches go through the library in one call (``n_trans``
  only bounds the fine-grid workspace; default: everything at once);
* sensitivity maps, density and sorted points always stay resident in HBM
  (``smaps_cached`` and ``async_transfer`` are accepted and ignored);
* host (numpy / torch-cpu) inputs are copied to the device and results are
  returned in the caller's array kind, like ``with_numpy_cupy``
  (``_array_compat.py:183-222``).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _capi
from ._array import from_device, kind_of, proper_trajectory, to_device
from .base import FourierOperatorBase, _ToggleGradPlanMixin, power_method, random_start

DEFAULT_WORKSPACE_BYTES = 24 << 30  # fine-grid workspace budget per operator


def _available():
    try:
        return _capi.LIB_PATH.exists() and torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


class RawHipPlan:
    """Type-2 plan plus the optional +1-sign gradient plan (``cufinufft.py:51-137``)."""

    def __init__(self, samples, shape, n_trans, eps, **opts):
        self.shape = tuple(shape)
        self.n_trans = n_trans
        self.eps = float(eps)
        self._opts = opts
        self.plan = _capi.RawPlan(shape, n_trans=n_trans, eps=eps, isign=-1, **opts)
        self.plan.setpts(samples)
        self.grad_plan = None

    def make_grad_plan(self, samples):
        self.grad_plan = _capi.RawPlan(self.shape, n_trans=self.n_trans, eps=self.eps, isign=1,
                                       **self._opts)
        self.grad_plan.setpts(samples)

    def set_pts(self, samples):
        self.plan.setpts(samples)
        if self.grad_plan is not None:
            self.grad_plan.setpts(samples)

    def toggle_grad_traj(self):
        if self.grad_plan is None:
            raise ValueError("gradient plan has not been initialised")
        self.plan, self.grad_plan = self.grad_plan, self.plan


class MRIHipNUFFT(FourierOperatorBase, _ToggleGradPlanMixin):
    """MRI NUFFT operator on MI355X (gfx950) through libhipnufft.

    Parameters follow ``MRICufiNUFFT`` (``cufinufft.py:202-216``): samples
    (M, d) or (Nc, Ns, d) in [-0.5, 0.5) or radians; shape; density (False,
    True / "pipe", dict, array or callable); n_coils; n_batchs; smaps;
    smaps_cached; verbose; squeeze_dims (default False); n_trans (None: all
    B*C transforms per launch within a 24 GiB workspace budget);
    async_transfer; and kwargs eps (1e-6), upsampfac (2.0), gpu_device_id,
    tile, spread_group, interp_group, max_subproblem, spread_max,
    spread_method ("auto", "cells", "blocks", "columns", "matrix"),
    interp_span (bins along the last axis per interpolation work item), and
    autotune (time
    candidate plan options on this trajectory and keep the fastest; explicit
    options win).
    """

    backend = "hipnufft"
    available = _available()
    autograd_available = True

    def __init__(self, samples, shape, density=False, n_coils=1, n_batchs=1, smaps=None,
                 smaps_cached=False, verbose=False, squeeze_dims=False, n_trans=None,
                 async_transfer=False, **kwargs):
        super().__init__()
        dev_id = int(kwargs.pop("gpu_device_id", 0))
        self.device = torch.device("cuda", dev_id)
        self.shape = shape
        self.n_batchs = n_batchs
        self.n_coils = n_coils
        self.squeeze_dims = squeeze_dims
        self.smaps_cached = True
        self.async_transfer = async_transfer
        self.verbose = verbose
        self._eps = float(kwargs.pop("eps", 1e-6))
        # options given by the caller (not None / 0 / "auto") are kept over
        # the autotuner's choice
        explicit = {k for k in ("tile", "spread_group", "interp_group", "max_subproblem",
                                "spread_max", "spread_method", "interp_span")
                    if kwargs.get(k) not in (None, 0, "auto")}
        self._plan_opts = dict(
            upsampfac=float(kwargs.pop("upsampfac", 2.0)),
            spreadinterponly=bool(kwargs.pop("spreadinterponly", False)),
            tile=kwargs.pop("tile", None),
            spread_group=int(kwargs.pop("spread_group", 0)),
            interp_group=int(kwargs.pop("interp_group", 0)),
            max_subproblem=int(kwargs.pop("max_subproblem", 0)),
            spread_max=int(kwargs.pop("spread_max", 0)),
            spread_method=kwargs.pop("spread_method", 0),
            interp_span=int(kwargs.pop("interp_span", 0)),
        )
        do_tune = bool(kwargs.pop("autotune", False))
        for k in ("gpu_method", "gpu_sort", "gpu_kerevalmeth", "gpu_spreadinterponly"):
            kwargs.pop(k, None)
        if kwargs:
            raise ValueError(f"unknown hipnufft options: {sorted(kwargs)}")

        with torch.cuda.device(self.device):
            traj = proper_trajectory(samples, normalize="pi")
            self._samples = to_device(traj, torch.float32, self.device)
            self.dtype = np.float32
            self._explicit_n_trans = n_trans is not None
            self.n_trans = self._choose_n_trans(n_trans)
            self.tuning = None
            if do_tune and not self._plan_opts["spreadinterponly"]:
                from .autotune import autotune, plan_options

                self.tuning = autotune(self._samples, self.shape, n_trans=self.n_trans,
                                       eps=self._eps, upsampfac=self._plan_opts["upsampfac"])
                for k, v in plan_options(self.tuning).items():
                    if k not in explicit:
                        self._plan_opts[k] = v
            self.raw_op = RawHipPlan(self._samples, self.shape, self.n_trans, self._eps,
                                     **self._plan_opts)
            if isinstance(density, (np.ndarray, torch.Tensor)):
                self.density = density
            else:
                self.compute_density(density)
            self.compute_smaps(smaps)
            self._check_n_trans()
            # load the device RNG of the power method's start vector now, not
            # inside the first reconstruction (its first use costs ~0.2 s)
            random_start((1,), self.device, 0)

    # ------------------------------------------------------------------ setup
    def _choose_n_trans(self, n_trans):
        BC = self.n_batchs * self.n_coils
        if n_trans is not None:
            if n_trans < 1:
                raise ValueError("n_trans must be >= 1")
            return int(n_trans)
        if self._plan_opts["spreadinterponly"]:
            return 1
        nf = [2 * s for s in self.shape]
        grid_bytes = 8 * int(np.prod(nf)) * max(1, int(self._plan_opts["upsampfac"] / 2.0) ** 2)
        return int(max(1, min(BC, DEFAULT_WORKSPACE_BYTES // max(grid_bytes, 1))))

    def _check_n_trans(self):
        if not self._explicit_n_trans:
            return
        if (self.n_batchs * self.n_coils) % self.n_trans != 0:
            raise ValueError(
                f"n_trans={self.n_trans} must divide n_batchs*n_coils={self.n_batchs * self.n_coils}")
        if self.uses_sense and self.n_coils % self.n_trans != 0:
            raise ValueError(f"n_trans={self.n_trans} must divide n_coils={self.n_coils}")

    @property
    def smaps(self):
        return self._smaps

    @smaps.setter
    def smaps(self, new_smaps):
        if new_smaps is None:
            self._smaps = None
            return
        if not isinstance(new_smaps, (np.ndarray, torch.Tensor)):
            raise ValueError("Smaps should be an array")
        C = new_smaps.shape[0]
        if tuple(new_smaps.shape[1:]) != self.shape:
            raise ValueError("Smaps shape does not match image shape.")
        if C != self.n_coils:
            self.log.warning("n_coils updated via smaps.")
            self.n_coils = C
        self._smaps = to_device(new_smaps, torch.complex64, self.device)

    @property
    def density(self):
        return self._density

    @density.setter
    def density(self, new_density):
        self._toeplitz_kernel = None  # the Toeplitz kernel holds the weights
        if new_density is None:
            self._density = None
            return
        d = new_density
        if isinstance(d, torch.Tensor) and d.is_complex():
            d = d.real
        elif isinstance(d, np.ndarray) and np.iscomplexobj(d):
            d = d.real
        d = to_device(d, torch.float32, self.device).reshape(-1)
        if d.numel() != self.n_samples:
            raise ValueError("Density and samples should have the same length")
        self._density = d

    @property
    def eps(self):
        return self.raw_op.eps

    def update_samples(self, new_samples, *, unsafe=False):
        """Re-bin the samples on the device and recompute the density (``cufinufft.py:317-347``)."""
        with torch.cuda.device(self.device):
            if unsafe:
                traj = new_samples
            else:
                traj = to_device(proper_trajectory(new_samples, normalize="pi"), torch.float32,
                                 self.device)
            if traj.ndim != 2 or traj.shape[1] != self.ndim:
                raise ValueError(f"samples should have shape (M, {self.ndim}), "
                                 f"got {tuple(traj.shape)}")
            # the new samples are committed only once the device sort has
            # succeeded; a failed sort leaves the plan without points, so
            # every transform raises until the next successful update
            self.raw_op.set_pts(traj)
            self._samples = traj
            self._toeplitz_kernel = None
            if self._density_method is not None:
                self.compute_density(self._density_method)

    def _make_plan_grad(self, **kwargs):
        self.raw_op.make_grad_plan(self._samples)

    # ------------------------------------------------------------- transforms
    def _img_in(self, data, ic):
        x = to_device(data, torch.complex64, self.device)
        return x.reshape(self.n_batchs, ic, *self.shape)

    def op(self, data, ksp=None):
        """Forward ``y[b, c] = F(S_c x_b) / norm`` -> (B, C, M) (squeezed if asked)."""
        self.check_shape(image=data, ksp=ksp)
        kind = kind_of(data)
        B, C, M = self.n_batchs, self.n_coils, self.n_samples
        with torch.cuda.device(self.device):
            x = self._img_in(data, 1 if self.uses_sense else C)
            out = torch.empty((B, C, M), dtype=torch.complex64, device=self.device)
            self.raw_op.plan.sense_op(x, self._smaps, out, B, C, self.inv_norm_factor)
        if ksp is not None and isinstance(ksp, torch.Tensor) and ksp.is_cuda:
            ksp.reshape(B, C, M).copy_(out)
        return from_device(self._safe_squeeze(out), kind)

    def adj_op(self, coeffs, img=None):
        """Adjoint ``x_b = sum_c conj(S_c) F^H(d * y[b, c]) / norm``."""
        self.check_shape(image=img, ksp=coeffs)
        kind = kind_of(coeffs)
        B, C, M = self.n_batchs, self.n_coils, self.n_samples
        with torch.cuda.device(self.device):
            y = to_device(coeffs, torch.complex64, self.device).reshape(B, C, M)
            ic = 1 if self.uses_sense else C
            out = torch.empty((B, ic, *self.shape), dtype=torch.complex64, device=self.device)
            if self._plan_opts["spreadinterponly"]:
                self.raw_op.plan.sense_adj(y, None, self._density, out, B, C, 1.0)
                out *= self.inv_norm_factor
            else:
                self.raw_op.plan.sense_adj(y, self._smaps, self._density, out, B, C,
                                           self.inv_norm_factor)
        return from_device(self._safe_squeeze(out), kind)

    def data_consistency(self, image_data, obs_data):
        """Fused ``A^H D (A x - y)`` (``cufinufft.py:992-1121``)."""
        obs_kind = kind_of(obs_data)
        self.check_shape(image=image_data, ksp=obs_data)
        B, C, M = self.n_batchs, self.n_coils, self.n_samples
        ic = 1 if self.uses_sense else C
        with torch.cuda.device(self.device):
            x = self._img_in(image_data, ic)
            y = to_device(obs_data, torch.complex64, self.device).reshape(B, C, M)
            out = torch.empty((B, ic, *self.shape), dtype=torch.complex64, device=self.device)
            self.raw_op.plan.sense_dc(x, self._smaps, y, self._density, out, B, C,
                                      self.inv_norm_factor)
        out = self._safe_squeeze(out)
        if obs_kind == "numpy" and kind_of(image_data) == "numpy":
            return from_device(out, "numpy")
        if kind_of(image_data) == "torch-cuda":
            return out
        return from_device(out, kind_of(image_data))

    def _op(self, image, coeffs):
        """Raw type 2 of ``image`` (T, *shape) into ``coeffs`` (T, M): no norm, no smaps."""
        T = image.shape[0] if image.ndim > self.ndim else 1
        x = image.reshape(T, *self.shape).resolve_conj().resolve_neg().contiguous()
        self.raw_op.plan.sense_op(x, None, coeffs, T, 1, 1.0)
        return coeffs

    def _adj_op(self, coeffs, image):
        """Raw type 1 of ``coeffs`` (T, M) into ``image`` (T, *shape): no norm, no density."""
        T = coeffs.reshape(-1, self.n_samples).shape[0]
        self.raw_op.plan.sense_adj(coeffs.resolve_conj().resolve_neg().contiguous(), None, None,
                                   image, T, 1, 1.0)
        return image

    # ---------------------------------------------------------------- Gram
    def compute_toeplitz_kernel(self):
        """Build the Toeplitz form of ``adj_op(op(.))`` (density included):
        one type-1 NUFFT of the weights onto a 2N grid and one FFT
        (``base.py:400-405``, ``toeplitz.py:35-269``)."""
        with torch.cuda.device(self.device):
            self._toeplitz_kernel = _capi.RawGram(
                self.shape, self._samples, self._density, eps=self._eps, isign=-1,
                scale=self.inv_norm_factor ** 2, n_trans=self.n_trans)
        return self._toeplitz_kernel

    def gram_op(self, data, toeplitz=True):
        """``A^H D A x`` -- by the Toeplitz kernel (no spreading or
        interpolation) or, with ``toeplitz=False``, as ``adj_op(op(x))``
        (``base.py:316-342``, ``toeplitz.py:376-468``)."""
        self.check_shape(image=data)
        if not toeplitz:
            return self.adj_op(self.op(data))
        kind = kind_of(data)
        B, C = self.n_batchs, self.n_coils
        ic = 1 if self.uses_sense else C
        with torch.cuda.device(self.device):
            if getattr(self, "_toeplitz_kernel", None) is None:
                self.compute_toeplitz_kernel()
            x = self._img_in(data, ic)
            out = torch.empty((B, ic, *self.shape), dtype=torch.complex64, device=self.device)
            self._toeplitz_kernel.apply(x, self._smaps, out, B, C)
        return from_device(self._safe_squeeze(out), kind)

    # --------------------------------------------------------------- extras
    def get_lipschitz_cst(self, max_iter=10, seed=None, toeplitz=False, **kwargs):
        """Power method on A^H D A without coils (``cufinufft.py:1239-1279``).

        The start vector is unseeded random as in the reference, drawn on the
        device; ``seed`` makes it reproducible (used by the serving loop so
        that a request gives the same result alone and among others).
        ``toeplitz=True`` iterates the Toeplitz form of the same operator
        (FFTs only, built if missing): equal to the NUFFT form at the NUFFT
        tolerance, and much cheaper for large trajectories, whose one-coil
        spreading is bound by staging the list entries."""
        n_coils, n_batchs, smaps, sq = self.n_coils, self.n_batchs, self._smaps, self.squeeze_dims
        self._smaps = None
        self.n_coils = 1
        self.n_batchs = 1
        self.squeeze_dims = True
        try:
            xt = random_start(self.shape, self.device, seed)
            A = self.gram_op if toeplitz else self
            L, _ = power_method(max_iter, A,
                                norm_func=lambda v: torch.linalg.vector_norm(v.reshape(-1)), x=xt)
        finally:
            self.n_coils = n_coils
            self.n_batchs = n_batchs
            self._smaps = smaps
            self.squeeze_dims = sq
        return float(L)

    @classmethod
    def pipe(cls, kspace_loc, volume_shape, max_iter=10, osf=2, normalize=True, **kwargs):
        """Pipe & Menon density compensation on the device (``cufinufft.py:1281-1331``).

        ``d <- d / |A_s A_s^H d|`` with spread/interp-only operators on the
        image grid (kernel chosen for ``upsampfac=osf``), then, if
        ``normalize``, ``d /= mean|A^H(d * A 1)|`` with the full NUFFT.
        Returns ``|d|`` (float32) in the caller's array kind.
        """
        if not cls.available:
            raise ValueError("hipnufft is not available, cannot estimate the density compensation")
        kind = kind_of(kspace_loc)
        kwargs = dict(kwargs)
        kwargs.pop("upsampfac", None)
        grid_op = cls(samples=kspace_loc, shape=volume_shape, upsampfac=osf,
                      spreadinterponly=True, **kwargs)
        dev = grid_op.device
        d = torch.ones(grid_op.n_samples, dtype=torch.complex64, device=dev)
        for _ in range(max_iter):
            d = d / torch.abs(grid_op.op(grid_op.adj_op(d)).reshape(-1))
        if normalize:
            test_op = cls(samples=kspace_loc, shape=volume_shape, **kwargs)
            test_im = torch.ones(volume_shape, dtype=torch.complex64, device=dev)
            recon = test_op.adj_op(d * test_op.op(test_im).reshape(-1))
            d = d / torch.mean(torch.abs(recon))
        return from_device(torch.abs(d.reshape(-1)), kind)

    def __repr__(self):
        return (
            "MRIHipNUFFT(\n"
            f"  shape: {self.shape}\n"
            f"  n_coils: {self.n_coils}\n"
            f"  n_samples: {self.n_samples}\n"
            f"  n_trans: {self.n_trans}\n"
            f"  n_batchs: {self.n_batchs}\n"
            f"  uses_density: {self.uses_density}\n"
            f"  uses_sense: {self.uses_sense}\n"
            f"  eps:{self.eps:.0e}\n"
            ")"
        )
