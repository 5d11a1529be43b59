"""Operator registry and base contract, mirroring mri-nufft's plugin interface.

The interface of ``src/mrinufft/operators/base.py`` (names, signatures,
error types and messages), implemented for this backend:

* ``check_backend`` / ``list_backends`` / ``get_operator``  (:45-179)
* ``FourierOperatorBase`` registry via ``__init_subclass__`` (:182-218),
  ``check_shape`` (:225-260), ``_safe_squeeze`` (:262-273), the shape /
  coil / batch / density / smaps properties (:692-835), ``norm_factor``
  (:837-850), ``compute_density`` (:576-624), ``compute_smaps`` (:407-448),
  ``get_lipschitz_cst`` (:626-665), ``pinv_solver`` (:667-690),
  ``make_autograd`` / ``with_autograd`` (:536-574, :863-877),
  ``make_linops`` (:450-490)
* ``power_method`` (:1155-1221) and ``_ToggleGradPlanMixin`` (:1224-1238)

The same registry names, error types and messages are used so that code
written against ``mrinufft.get_operator`` keeps working; when the real
``mrinufft`` package is importable the backend is also registered there
(see ``hipnufft.register_into_mrinufft``).
"""

from __future__ import annotations

import logging
import warnings
from contextlib import contextmanager
from functools import partial

import numpy as np
import torch

from ._array import kind_of

DTYPE_R2C = {"float32": "complex64", "float64": "complex128"}


def _registered(backend_name: str):
    """(available, class) of a registered backend, or None."""
    return FourierOperatorBase.interfaces.get(backend_name.lower())


def check_backend(backend_name: str):
    """Availability flag of a registered backend (ValueError if unknown)."""
    entry = _registered(backend_name)
    if entry is None:
        raise ValueError(f"unknown backend: '{backend_name.lower()}'")
    return entry[0]


def list_backends(available_only=False):
    """Registered backend names (only the usable ones with ``available_only``)."""
    names = list(FourierOperatorBase.interfaces)
    return [n for n in names if check_backend(n)] if available_only else names


def get_operator(backend_name: str, wrt_data: bool = False, wrt_traj: bool = False,
                 paired_batch: bool = False, *args, **kwargs):
    """The operator class of ``backend_name``; an instance when constructor
    arguments are given; wrapped for autograd when ``wrt_data`` /
    ``wrt_traj`` is set (an instance is wrapped now, a class when it is
    built).  ValueError for an unknown or unavailable backend."""
    name = backend_name.lower()
    entry = _registered(name)
    if entry is None:
        raise ValueError(f"backend {name} does not exist")
    available, klass = entry
    if not available:
        raise ValueError(
            f"backend {name} found, but dependencies are not met."
            f" ``pip install mri-nufft[{name}]`` may solve the issue."
        )
    target = klass(*args, **kwargs) if (args or kwargs) else klass
    if not (wrt_data or wrt_traj):
        return target
    if isinstance(target, partial):
        raise ValueError("Cannot create autograd of a partial operator.")
    if isinstance(target, FourierOperatorBase):
        return target.make_autograd(wrt_data=wrt_data, wrt_traj=wrt_traj,
                                    paired_batch=paired_batch)
    return partial(klass.with_autograd, wrt_data, wrt_traj, paired_batch)


def _positive_count(value, what):
    """``value`` as an int when it is a whole number >= 1, else ValueError."""
    if value < 1 or int(value) != value:
        raise ValueError(f"{what} should be a positive integer, {type(value)}")
    return int(value)


class FourierOperatorBase:
    """Base Fourier operator: registry and the shared MRI contract."""

    interfaces: dict = {}
    autograd_available = False
    backend: str
    available: bool

    def __init__(self):
        if not self.available:
            raise RuntimeError(f"'{self.backend}' backend is not available.")
        self._smaps = None
        self._density = None
        self._n_coils = 1
        self._n_batchs = 1
        self.n_trans = 1
        self.squeeze_dims = False
        self._density_method = None
        self._grad_wrt_data = False
        self._grad_wrt_traj = False
        self._grad_wrt_field_map = False

    def __init_subclass__(cls):
        """Defining a subclass with a ``backend`` name registers it, with its
        ``available`` flag (called once if it is a function)."""
        super().__init_subclass__()
        name = getattr(cls, "backend", None)
        if name:
            flag = getattr(cls, "available", True)
            cls.interfaces[name] = (flag() if callable(flag) else flag, cls)

    @property
    def log(self):
        return logging.getLogger(type(self).__module__)

    # ------------------------------------------------------------------ shapes
    def check_shape(self, *, image=None, ksp=None):
        """ValueError when the image's trailing axes differ from the operator
        shape or the k-space's last axis from the number of samples."""
        if image is None and ksp is None:
            raise ValueError("Nothing to check, provides image or ksp arguments")
        got = None if image is None else tuple(image.shape[-self.ndim:])
        if got is not None and got != self.shape:
            raise ValueError(f"Image shape {got} is not compatible "
                             f"with the operator shape {self.shape}")
        if ksp is not None and ksp.shape[-1] != self.n_samples:
            raise ValueError(f"Kspace samples {ksp.shape[-1]} is not compatible "
                             f"with the operator samples {self.n_samples}")

    def _safe_squeeze(self, arr):
        """Squeeze the coil axis, then the batch axis, when they are singleton."""
        if self.squeeze_dims:
            for ax in (1, 0):
                if arr.ndim > ax and arr.shape[ax] == 1:
                    arr = arr.squeeze(ax) if isinstance(arr, torch.Tensor) else arr.squeeze(axis=ax)
        return arr

    @property
    def uses_sense(self):
        return self._smaps is not None

    @property
    def uses_density(self):
        return self.density is not None

    @property
    def ndim(self):
        return len(self._shape)

    @property
    def shape(self):
        return self._shape

    @shape.setter
    def shape(self, shape):
        self._shape = tuple(int(i) for i in shape)

    @property
    def n_coils(self):
        return self._n_coils

    @n_coils.setter
    def n_coils(self, n_coils):
        self._n_coils = _positive_count(n_coils, "n_coils")

    @property
    def n_batchs(self):
        return self._n_batchs

    @n_batchs.setter
    def n_batchs(self, n_batchs):
        self._n_batchs = _positive_count(n_batchs, "n_batchs")

    @property
    def img_full_shape(self):
        return (self.n_batchs, (1 if self.uses_sense else self.n_coils)) + self.shape

    @property
    def ksp_full_shape(self):
        return (self.n_batchs, self.n_coils, self.n_samples)

    @property
    def density(self):
        return self._density

    @density.setter
    def density(self, new_density):
        if new_density is None:
            self._density = None
        elif len(new_density) != self.n_samples:
            raise ValueError("Density and samples should have the same length")
        else:
            self._density = new_density

    @property
    def dtype(self):
        return self._dtype

    @dtype.setter
    def dtype(self, new_dtype):
        self._dtype = np.dtype(new_dtype)

    @property
    def cpx_dtype(self):
        return np.dtype(DTYPE_R2C[str(self.dtype)])

    @property
    def samples(self):
        return self._samples

    @samples.setter
    def samples(self, new_samples):
        self.update_samples(new_samples)

    def update_samples(self, new_samples, *, unsafe=False):
        self._samples = new_samples

    @property
    def n_samples(self):
        return self._samples.shape[0]

    @property
    def norm_factor(self):
        """``sqrt(prod(shape) * 2**ndim)`` (``base.py:837-840``)."""
        return float(np.sqrt(np.prod(self.shape) * (2 ** len(self.shape))))

    @property
    def inv_norm_factor(self):
        return 1.0 / self.norm_factor

    # -------------------------------------------------------------- methods
    def data_consistency(self, image_data, obs_data):
        """Naive ``adj_op(op(x) - y)``; backends override with a fused version."""
        return self.adj_op(self.op(image_data) - obs_data)

    def gram_op(self, data, toeplitz: bool = False):
        """``adj_op(op(x))``; the Toeplitz embedding is not provided by this backend."""
        self.check_shape(image=data)
        if toeplitz:
            raise NotImplementedError("Toeplitz Gram operator is not implemented by hipnufft")
        return self.adj_op(self.op(data))

    def _density_estimator(self, method):
        """(function, keyword arguments) of a density request: True or
        "pipe" (Pipe on this backend), a registered name, a dict with a
        "name" and that method's arguments, or a callable."""
        from .density import get_density

        if method is True or method == "pipe":
            return get_density("pipe"), {"backend": self.backend}
        if isinstance(method, dict) and isinstance(method.get("name"), str):
            opts = {k: v for k, v in method.items() if k != "name"}
            if method["name"] == "pipe":
                opts.setdefault("backend", self.backend)
            return get_density(method["name"]), opts
        if isinstance(method, str):
            return get_density(method), {}
        if callable(method):
            return method, {}
        raise ValueError(f"Unknown density method: {method}")

    def compute_density(self, method=None):
        """Set the density compensation from an array, or from an estimator
        (see ``_density_estimator``) run on the samples; None / False
        removes it.  The first estimator used is kept to re-run on
        ``update_samples``."""
        if method is None or method is False:
            self.density = None
        elif isinstance(method, (np.ndarray, torch.Tensor)):
            self.density = method
        else:
            estimate, opts = self._density_estimator(method)
            self.density = estimate(self.samples, self.shape, **opts)
            if self._density_method is None:
                self._density_method = partial(estimate, **opts)

    def compute_smaps(self, method=None):
        """Set the sensitivity maps (array or callable(samples, shape, ...))."""
        if method is None:
            self.smaps = None
            return
        if isinstance(method, (np.ndarray, torch.Tensor)):
            self.smaps = method.reshape(self.n_coils if method.shape[0] == self.n_coils
                                        else method.shape[0], *self.shape)
            return
        if callable(method):
            smaps = method(self.samples, self.shape, density=self.density, backend=self.backend)
            self.smaps = smaps.reshape(-1, *self.shape)
            return
        raise ValueError(f"Unknown smaps method: {method}")

    def get_lipschitz_cst(self, max_iter=10):
        """Spectral radius of A^H A without coils (power method)."""
        n_coils, n_batchs, smaps, sq = self.n_coils, self.n_batchs, self.smaps, self.squeeze_dims
        self.smaps = None
        self.n_coils = 1
        self.n_batchs = 1
        self.squeeze_dims = True
        try:
            lipschitz_cst, _ = power_method(max_iter, self)
        finally:
            self.n_coils = n_coils
            self.n_batchs = n_batchs
            self.smaps = smaps
            self.squeeze_dims = sq
        return lipschitz_cst

    def with_off_resonance_correction(self, readout_time, b0_map=None, r2star_map=None, mask=None,
                                      interpolator="svd"):
        """A new operator with off-resonance correction (``base.py:385-398``)."""
        from .off_resonance import MRIFourierCorrected

        return MRIFourierCorrected(self, b0_map, readout_time, r2star_map, mask, interpolator)

    def pinv_solver(self, kspace_data, optim="lsqr", **kwargs):
        """Least-squares reconstruction with a registered optimizer."""
        from .optim import get_optimizer

        return get_optimizer(optim)(operator=self, kspace_data=kspace_data, **kwargs)

    def make_autograd(self, *, wrt_data=True, wrt_traj=False, paired_batch=False):
        if not self.autograd_available:
            raise ValueError("Backend does not support auto-differentiation.")
        from .autodiff import MRINufftAutoGrad

        return MRINufftAutoGrad(self, wrt_data=wrt_data, wrt_traj=wrt_traj,
                                paired_batch=paired_batch)

    @classmethod
    def with_autograd(cls, wrt_data=True, wrt_traj=False, paired_batch=False, *args, **kwargs):
        return cls(*args, **kwargs).make_autograd(wrt_data=wrt_data, wrt_traj=wrt_traj,
                                                  paired_batch=paired_batch)

    def make_linops(self, *, cupy: bool = False):
        """scipy LinearOperator over flattened host arrays (``base.py:450-490``)."""
        if cupy:
            raise ValueError("cupy is not available")
        from scipy.sparse.linalg import LinearOperator

        ic = 1 if self.uses_sense else self.n_coils
        linop = LinearOperator(
            shape=(self.n_batchs * self.n_coils * self.n_samples,
                   self.n_batchs * ic * int(np.prod(self.shape))),
            matvec=lambda x: np.asarray(self.op(x.reshape(self.n_batchs, ic, *self.shape))).ravel(),
            rmatvec=lambda x: np.asarray(
                self.adj_op(x.reshape(self.n_batchs, self.n_coils, self.n_samples))).ravel(),
            dtype=self.cpx_dtype,
        )
        linop._nufft = self
        return linop

    def __repr__(self):
        return (f"{self.__class__.__name__}(\n  shape: {self.shape}\n  n_coils: {self.n_coils}\n"
                f"  n_samples: {self.n_samples}\n  uses_sense: {self.uses_sense}\n)")


def random_start(shape, device, seed=None):
    """``1j * U[0, 1) + U[0, 1)`` complex64 start vector of the power method,
    drawn on ``device`` (the reference draws the same distribution with numpy
    on the host, ``base.py:1155-1221``); seeded when ``seed`` is given."""
    gen = torch.Generator(device=device)
    if seed is None:
        gen.seed()
    else:
        gen.manual_seed(int(seed))
    im = torch.rand(tuple(shape), generator=gen, device=device, dtype=torch.float32)
    re = torch.rand(tuple(shape), generator=gen, device=device, dtype=torch.float32)
    return torch.complex(re, im)


def _l2(v):
    if isinstance(v, torch.Tensor):
        return torch.linalg.vector_norm(v.reshape(-1))
    return np.linalg.norm(v)


def power_method(max_iter, operator, norm_func=None, x=None):
    """Largest eigenvalue of A^H A by power iterations (interface and stopping
    rule of ``operators/base.py:1155-1221``).

    ``operator`` is an operator (A^H A = adj_op(op(.))) or a callable applying
    A^H A.  Starting from ``x`` (required for a callable; otherwise
    ``U[0, 1)`` on the host, complex), each step applies A^H A to the unit
    vector and takes the norm of the result as the estimate.  The iterations
    stop when two successive norms differ by less than 1e-6 (the first one
    against the start vector's norm), after at most ``max_iter`` steps; a
    warning is issued when all ``max_iter`` steps ran.  Returns
    (estimate, unit eigenvector estimate); the estimate is a python float
    when the start vector was drawn here.  Works on numpy arrays and torch
    tensors.
    """
    if callable(operator) and not isinstance(operator, FourierOperatorBase):
        gram = operator
    else:
        def gram(v):
            return operator.adj_op(operator.op(v))
    norm = norm_func or _l2
    drawn = x is None
    if drawn:
        if not isinstance(operator, FourierOperatorBase):
            raise ValueError("If operator is not a FourierOperatorBase, an initial x should "
                             "be provided.")
        x = np.random.random(operator.shape).astype(operator.cpx_dtype)
    last = norm(x)
    vec = x / last
    estimate = last
    steps = 0
    while steps < max_iter:
        steps += 1
        image = gram(vec)
        estimate = norm(image)
        vec = image / estimate
        if abs(float(last) - float(estimate)) < 1e-6:
            break
        last = estimate
    if max_iter and steps == max_iter:
        warnings.warn("Lipschitz constant did not converge")
    return (float(estimate) if drawn else estimate), vec


class _ToggleGradPlanMixin:
    """Swaps the operator's type-2 plan with its +1-sign plan (and its coil
    maps with their conjugates) for the trajectory-gradient products
    (``base.py:1224-1238``)."""

    @contextmanager
    def grad_traj_plan(self):
        """Run the body on the gradient plan; the swap is undone on exit."""
        self.toggle_grad_traj()
        try:
            yield
        finally:
            self.toggle_grad_traj()

    def toggle_grad_traj(self):
        if self.uses_sense:
            self.smaps = self.smaps.conj()
        self.raw_op.toggle_grad_traj()


def array_kind(x):
    return kind_of(x)
