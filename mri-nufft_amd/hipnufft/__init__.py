"""hipnufft: an MI355X-native NUFFT backend with mri-nufft's operator interface.

``get_operator("hipnufft", samples, shape, density=..., n_coils=..., smaps=...)``
returns an operator with mri-nufft's ``op`` / ``adj_op`` / ``data_consistency``
/ ``pinv_solver`` / ``get_lipschitz_cst`` surface; the arithmetic runs in
``libhipnufft.so`` (hand-written gfx950 HIP kernels) reached through a ctypes
C ABI on PyTorch-ROCm device tensors.  There is no CPU path: on a machine
without the library or a GPU the backend is registered as unavailable and
every transform raises.
"""

from .base import FourierOperatorBase, check_backend, get_operator, list_backends, power_method
from .density import get_density, pipe
from .field_map import (get_complex_fieldmap_rad, get_orc_factorization, make_b0map,
                        make_t2smap)
from .off_resonance import MRIFourierCorrected
from .operator import MRIHipNUFFT
from .optim import cg, get_optimizer, lsmr, lsqr
from .smaps import get_smaps, low_frequency
from .stacked import MRIStackedHipNUFFT, MRIStackedNUFFT, stacked2traj3d, traj3d2stacked

__all__ = [
    "FourierOperatorBase",
    "MRIFourierCorrected",
    "MRIHipNUFFT",
    "MRIStackedHipNUFFT",
    "MRIStackedNUFFT",
    "cg",
    "check_backend",
    "get_complex_fieldmap_rad",
    "get_density",
    "get_operator",
    "get_orc_factorization",
    "get_optimizer",
    "get_smaps",
    "low_frequency",
    "list_backends",
    "lsmr",
    "lsqr",
    "make_b0map",
    "make_t2smap",
    "pipe",
    "power_method",
    "register_into_mrinufft",
    "stacked2traj3d",
    "traj3d2stacked",
]


def register_into_mrinufft():
    """Register the backend in an installed ``mrinufft`` (no-op if absent).

    Defines ``mrinufft.operators.interfaces.hipnufft.MRIHipNUFFT`` as a
    subclass of mrinufft's ``FourierOperatorBase`` so that
    ``mrinufft.get_operator("hipnufft", ...)`` returns this backend.
    """
    try:
        import mrinufft.operators.base as mb
    except ImportError:
        return False
    registry = mb.FourierOperatorBase.__dict__["interfaces"]
    if "hipnufft" in registry:
        return True
    own = FourierOperatorBase.interfaces.get("hipnufft")
    cls = type("MRIHipNUFFT", (MRIHipNUFFT, mb.FourierOperatorBase),
               {"backend": "hipnufft", "available": MRIHipNUFFT.available,
                "__module__": "mrinufft.operators.interfaces.hipnufft"})
    if own is not None:  # keep hipnufft's own registry pointing at the plain class
        FourierOperatorBase.interfaces["hipnufft"] = own
    # hipnufft's base comes first in the MRO, so ``cls.interfaces`` is
    # hipnufft's own dict: register in mri-nufft's registry explicitly
    registry["hipnufft"] = (MRIHipNUFFT.available, cls)
    return True
