"""Density-compensation registry and the NUFFT-based Pipe dispatcher.

Restates ``src/mrinufft/density/utils.py:11-33`` (``register_density``,
``get_density``, ``flat_traj``) and ``density/nufft_based.py:6-28``
(``pipe``).  The geometric methods (voronoi, cell_count) of the reference are
CPU/scipy code outside the GPU path and are not provided.
"""

from __future__ import annotations

from functools import wraps

from ._array import proper_trajectory
from ._registry import MethodRegister

register_density = MethodRegister("density")
get_density = register_density.make_getter()


def flat_traj(normalize="unit"):
    """Decorator: flatten / normalise the trajectory before calling."""

    def decorator(func):
        @wraps(func)
        def wrapper(*args, **kwargs):
            args = list(args)
            args[0] = proper_trajectory(args[0], normalize=normalize)
            return func(*args, **kwargs)

        return wrapper

    if callable(normalize):
        func = normalize
        normalize = "unit"
        return decorator(func)
    return decorator


@register_density
@flat_traj
def pipe(traj, shape, backend="hipnufft", **kwargs):
    """Pipe & Menon density compensation through the backend's ``pipe``."""
    from .base import get_operator

    nufft_class = get_operator(backend)
    if hasattr(nufft_class, "pipe"):
        return nufft_class.pipe(traj, shape, **kwargs)
    raise ValueError("backend does not have pipe iterations method.")
