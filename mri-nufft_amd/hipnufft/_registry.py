"""Named tables of interchangeable functions (optimizers, density and
coil-map estimators, off-resonance factorizations).

Interface of mri-nufft's ``MethodRegister`` (``src/mrinufft/_utils.py:162-228``):
an instance decorates functions of one kind, ``make_getter()`` returns
``get_<kind>(name, *args, **kwargs)`` (the function, or its result when
arguments are given; ValueError with the available names otherwise) and
``registry`` is the kind's name -> function table.  Tables are shared by
every instance created for the same kind.
"""

from __future__ import annotations

from typing import Callable

_TABLES: dict[str, dict[str, Callable]] = {}


class MethodRegister:
    """Register functions of one kind under a name (the function's own
    name unless one is given)."""

    def __init__(self, register_name: str):
        self.register_name = register_name
        _TABLES.setdefault(register_name, {})

    @property
    def registry(self) -> dict[str, Callable]:
        return _TABLES[self.register_name]

    def add(self, func: Callable, name: str | None = None) -> Callable:
        """Store ``func`` under ``name`` (default ``func.__name__``); returns it."""
        self.registry[name or func.__name__] = func
        return func

    def __call__(self, method_name=None):
        """``@register`` or ``@register("name")``."""
        if callable(method_name):
            return self.add(method_name)
        return lambda func: self.add(func, method_name)

    def make_getter(self) -> Callable:
        kind = self.register_name
        table = self.registry

        def getter(method_name, *args, **kwargs):
            func = table.get(method_name)
            if func is None:
                raise ValueError(f"Unknown {kind} method {method_name}. Available methods are \n"
                                 f"{list(table)}")
            return func(*args, **kwargs) if (args or kwargs) else func

        getter.__name__ = f"get_{kind}"
        getter.__doc__ = f"The {kind} method registered as ``method_name`` (called if arguments are given)."
        return getter
