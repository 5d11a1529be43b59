"""Host-side checks that need no GPU.

* the C-ABI library loads and exports every entry point declared in
  include/hipnufft.h (no compute call is made);
* registry semantics of get_operator / list_backends / check_backend
  (reference tests/operators/test_bindings.py:7-30, operators/base.py:45-179);
* proper_trajectory (reference _utils.py:58-100) and the method registries;
* trajectory generators against their closed forms;
* bench.py refuses more than one GPU.
"""

import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "hipnufft.h")).read()
    return sorted(set(re.findall(r"\b(hipnufft_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    from hipnufft import _capi

    lib = _capi.lib()
    declared = header_symbols()
    assert "hipnufft_plan_create" in declared and "hipnufft_sense_dc" in declared
    for name in declared:
        assert hasattr(lib, name), name
    # the ctypes table covers exactly the declared entry points
    assert sorted(_capi.EXPORTED) == declared


def test_opts_struct_layout():
    import ctypes

    from hipnufft import _capi

    o = _capi.Opts()
    assert _capi.lib().hipnufft_default_opts(ctypes.byref(o)) == 0
    assert o.upsampfac == 2.0 and list(o.tile) == [0, 0, 0] and o.spreadinterponly == 0


def test_registry_semantics():
    import hipnufft
    from hipnufft.base import FourierOperatorBase

    assert "hipnufft" in hipnufft.list_backends()
    with pytest.raises(ValueError):
        hipnufft.check_backend("nope")
    with pytest.raises(ValueError):
        hipnufft.get_operator("nope")
    avail = hipnufft.check_backend("hipnufft")
    if not avail:
        with pytest.raises(ValueError, match="dependencies are not met"):
            hipnufft.get_operator("hipnufft")
        assert "hipnufft" not in hipnufft.list_backends(available_only=True)

    class Fake(FourierOperatorBase):
        backend = "fake-test"
        available = False

    assert FourierOperatorBase.interfaces["fake-test"] == (False, Fake)
    with pytest.raises(RuntimeError):
        Fake()
    del FourierOperatorBase.interfaces["fake-test"]


def test_method_registries():
    from hipnufft import get_density, get_optimizer

    assert callable(get_optimizer("cg")) and callable(get_optimizer("lsqr"))
    assert callable(get_optimizer("lsmr")) and callable(get_density("pipe"))
    with pytest.raises(ValueError, match="Unknown optimizer method"):
        get_optimizer("sgd")


def test_proper_trajectory():
    import torch

    from hipnufft._array import proper_trajectory

    unit = np.array([[[0.25, -0.5], [0.1, 0.0]]], dtype=np.float32)
    out = proper_trajectory(unit, "pi")
    assert out.shape == (2, 2) and out.dtype == np.float32
    np.testing.assert_allclose(out, unit.reshape(-1, 2) * 2 * np.pi, rtol=1e-6)
    rad = np.array([[np.pi / 2, -3.0]])
    np.testing.assert_allclose(proper_trajectory(rad, "pi"), rad)
    np.testing.assert_allclose(proper_trajectory(rad, "unit"), rad / (2 * np.pi))
    t = torch.tensor([[0.25, 0.5]])
    np.testing.assert_allclose(proper_trajectory(t, "unit").numpy(), [[0.25, -0.5]])


def test_radial_closed_form():
    from hipnufft.trajectories import initialize_2D_radial

    Nc, Ns = 16, 33
    t = initialize_2D_radial(Nc, Ns)
    r = 0.5 * np.linspace(0, 1, Ns)
    th = 2 * np.pi * np.arange(Nc) / Nc
    ref = np.stack([np.cos(th)[:, None] * r, np.sin(th)[:, None] * r], -1)
    np.testing.assert_allclose(t, ref, atol=1e-13)
    t2 = initialize_2D_radial(8, 9, in_out=True)
    np.testing.assert_allclose(np.linalg.norm(t2, axis=-1).max(), 0.5)
    np.testing.assert_allclose(t2[:, 4], 0, atol=1e-15)


def test_spiral_closed_form():
    from hipnufft.trajectories import initialize_2D_spiral

    Nc, Ns, rev = 4, 65, 2
    t = initialize_2D_spiral(Nc, Ns, nb_revolutions=rev)
    a = 2 * np.pi * rev * np.linspace(0, 1, Ns)
    shot = a * np.exp(1j * a) / a.max() * 0.5
    for i in range(Nc):
        z = shot * np.exp(1j * 2 * np.pi * i / Nc)
        np.testing.assert_allclose(t[i, :, 0] + 1j * t[i, :, 1], z, atol=1e-13)


def test_phyllotaxis_closed_form():
    from hipnufft.trajectories import GOLDEN, initialize_3D_phyllotaxis_radial

    Nc, Ns = 64, 17
    t = initialize_3D_phyllotaxis_radial(Nc, Ns)
    z = np.linspace(-1, 1, Nc)
    phi = GOLDEN * np.arange(Nc)
    v = np.stack([np.cos(phi) * np.sqrt(1 - z ** 2), np.sin(phi) * np.sqrt(1 - z ** 2), z], -1)
    r = 0.5 * np.linspace(0, 1, Ns)
    ref = -r[None, :, None] * v[:, None, :]
    np.testing.assert_allclose(t, ref, atol=1e-12)


def test_stack_closed_form():
    from hipnufft.trajectories import initialize_2D_radial, stack

    base = initialize_2D_radial(4, 5)
    t = stack(base, 3)
    assert t.shape == (12, 5, 3)
    np.testing.assert_allclose(t[:4, :, :2], base)
    np.testing.assert_allclose(np.unique(t[:, :, 2]), [-1 / 3, 0, 1 / 3], atol=1e-15)


def test_smaps_phantom():
    from hipnufft.trajectories import ellipsoid_phantom, gaussian_smaps

    s = gaussian_smaps(5, (8, 9, 10), seed=3)
    np.testing.assert_allclose(np.sum(np.abs(s.astype(np.complex128)) ** 2, 0), 1, rtol=1e-5)
    p1 = ellipsoid_phantom((16, 16), seed=2)
    p2 = ellipsoid_phantom((16, 16), seed=2)
    np.testing.assert_array_equal(p1, p2)
    assert p1.dtype == np.complex64


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_register_into_mrinufft(tmp_path, monkeypatch):
    """A minimal stand-in package with mri-nufft's registry contract
    (operators/base.py:182-218: abstract op/adj_op/_op/_adj_op, a class-level
    ``interfaces`` dict filled by ``__init_subclass__``)."""
    pkg = tmp_path / "mrinufft" / "operators"
    pkg.mkdir(parents=True)
    (tmp_path / "mrinufft" / "__init__.py").write_text("")
    (pkg / "__init__.py").write_text("")
    (pkg / "base.py").write_text(
        "from abc import ABC, abstractmethod\n"
        "class FourierOperatorBase(ABC):\n"
        "    interfaces = {}\n"
        "    def __init_subclass__(cls):\n"
        "        super().__init_subclass__()\n"
        "        name = getattr(cls, 'backend', None)\n"
        "        if name:\n"
        "            cls.interfaces[name] = (getattr(cls, 'available', True), cls)\n"
        "    @abstractmethod\n"
        "    def op(self, x): ...\n"
        "    @abstractmethod\n"
        "    def adj_op(self, y): ...\n"
        "    @abstractmethod\n"
        "    def _op(self, x, y): ...\n"
        "    @abstractmethod\n"
        "    def _adj_op(self, y, x): ...\n")
    monkeypatch.syspath_prepend(str(tmp_path))
    for m in [m for m in sys.modules if m.startswith("mrinufft")]:
        monkeypatch.delitem(sys.modules, m)
    import hipnufft
    from hipnufft.base import FourierOperatorBase as OwnBase

    own_before = OwnBase.interfaces["hipnufft"]
    assert hipnufft.register_into_mrinufft() is True
    import mrinufft.operators.base as mb

    avail, cls = mb.FourierOperatorBase.interfaces["hipnufft"]
    assert issubclass(cls, mb.FourierOperatorBase) and issubclass(cls, hipnufft.MRIHipNUFFT)
    assert not getattr(cls, "__abstractmethods__", None)
    assert avail == hipnufft.MRIHipNUFFT.available
    assert OwnBase.interfaces["hipnufft"] == own_before
    assert hipnufft.register_into_mrinufft() is True  # idempotent


def test_autotune_candidates_respect_kernel_limits():
    from hipnufft.autotune import tile_candidates, tile_ok

    assert tile_ok((8, 8, 16), 7) and not tile_ok((8, 16, 16), 7)   # 512 segments
    assert not tile_ok((16, 16, 16), 7)
    assert tile_ok((32, 32), 7, interp_group=4) and not tile_ok((32, 32), 7, interp_group=6)
    for nd in (1, 2, 3):
        for w in (4, 7, 10):
            cands = tile_candidates(nd, w)
            assert cands and all(len(c) == nd and tile_ok(c, w) for c in cands)
            assert all(t & (t - 1) == 0 for c in cands for t in c)  # powers of two


def test_toeplitz_support_rule():
    from hipnufft.serving import toeplitz_supported

    assert toeplitz_supported((192, 192, 192)) and toeplitz_supported((48, 48))
    assert not toeplitz_supported((40, 56))      # 112 = 2^4 * 7
    assert not toeplitz_supported((1024, 1024))  # kernel build grid 4096 > 2730


def test_stack_layout_detection():
    """The server recognises a stack of one 2D pattern on Cartesian kz
    planes (in unit or radian coordinates) and nothing else."""
    import numpy as np

    from hipnufft import stacked2traj3d
    from hipnufft.serving import stack_layout

    s2 = np.random.default_rng(0).uniform(-0.5, 0.5, (50, 2)).astype(np.float32)
    zi = np.array([3, 0, 7, 5])
    t3 = stacked2traj3d(s2, zi, 8)
    got = stack_layout(t3, (16, 16, 8))
    assert got is not None
    np.testing.assert_array_equal(got[1], zi)
    np.testing.assert_allclose(got[0], s2, atol=1e-7)
    rad = (t3 * np.float32(2 * np.pi)).astype(np.float32)
    np.testing.assert_array_equal(stack_layout(rad, (16, 16, 8))[1], zi)
    assert stack_layout(t3[:, :2], (16, 16)) is None                 # 2D
    shifted = t3.copy()
    shifted[:, 2] += 0.5 / 8                                         # half-integer kz
    assert stack_layout(shifted, (16, 16, 8)) is None
    jitter = t3.copy()
    jitter[60, 0] += 1e-3                                            # plane 2 differs
    assert stack_layout(jitter, (16, 16, 8)) is None
    near = t3.astype(np.float64)
    near[60, 0] += 5e-8                                              # within 1e-7: a stack
    got = stack_layout(near, (16, 16, 8))
    assert got is not None and np.array_equal(got[1], zi)
    assert stack_layout(np.random.default_rng(1).uniform(-0.5, 0.5, (200, 3)), (8, 8, 8)) is None
