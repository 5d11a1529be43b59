"""Named method registries (``src/mrinufft/_utils.py:162-228``, ``MethodRegister``)."""

from __future__ import annotations

from collections import defaultdict


class MethodRegister:
    """Decorator registering functions of one kind under a name."""

    registry_global = defaultdict(dict)

    def __init__(self, register_name: str):
        self.register_name = register_name

    def __call__(self, method_name=None):
        def decorator(func):
            self.registry[method_name] = func
            return func

        if callable(method_name):
            func = method_name
            method_name = func.__name__
            return decorator(func)
        return decorator

    def make_getter(self):
        def getter(method_name, *args, **kwargs):
            try:
                method = self.registry[method_name]
            except KeyError as e:
                raise ValueError(
                    f"Unknown {self.register_name} method {method_name}."
                    " Available methods are \n"
                    f"{list(self.registry.keys())}"
                ) from e
            if args or kwargs:
                return method(*args, **kwargs)
            return method

        getter.__name__ = f"get_{self.register_name}"
        return getter

    @property
    def registry(self):
        return self.registry_global[self.register_name]
