"""The entry point that allocates nothing, ``hipnufft_dot_c64_rows_noalloc``,
called through the C ABI: each row equals ``hipnufft_dot_c64_rows`` bit for
bit (same kernels and summation order, its partials in the library's static
device array instead of a workspace), on two calls in a row, the second with
a smaller input, so stale partials of the first call cannot leak into the
second; sizes over the limit are rejected with an argument error."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _pair(rows, n, seed):
    g = torch.Generator(device=DEV).manual_seed(seed)
    a = torch.complex(torch.randn((rows, n), generator=g, device=DEV),
                      torch.randn((rows, n), generator=g, device=DEV))
    b = torch.complex(torch.randn((rows, n), generator=g, device=DEV),
                      torch.randn((rows, n), generator=g, device=DEV))
    return a, b


def _noalloc(a, b, conj, out):
    from hipnufft import _capi

    return _capi.lib().hipnufft_dot_c64_rows_noalloc(
        _capi.ptr(a), _capi.ptr(b), a.shape[0], a.shape[1], 1 if conj else 0, _capi.ptr(out),
        _capi.current_stream())


def test_noalloc_rows_equal_allocating_rows_twice():
    from hipnufft import _capi

    for conj in (True, False):
        # large first call (every row and every partial block in use), then a
        # smaller one: fewer rows and fewer partial blocks per row
        for rows, n, seed in ((_capi.NOALLOC_DOT_ROWS, 1 << 20, 1), (3, 700, 2)):
            a, b = _pair(rows, n, seed)
            want = _capi.dot_c64_rows(a, b, conj=conj)
            out = torch.full((rows, 2), float("nan"), dtype=torch.float64, device=DEV)
            assert _noalloc(a, b, conj, out) == 0
            torch.cuda.synchronize()
            assert torch.equal(out, want), (conj, rows, n)
            # and against an FP64 reference of the same dots
            a64, b64 = a.to(torch.complex128), b.to(torch.complex128)
            ref = torch.sum((a64.conj() if conj else a64) * b64, dim=1)
            np.testing.assert_allclose(out[:, 0].cpu().numpy(), ref.real.cpu().numpy(),
                                       rtol=1e-9, atol=1e-9 * float(ref.abs().max()))
            np.testing.assert_allclose(out[:, 1].cpu().numpy(), ref.imag.cpu().numpy(),
                                       rtol=1e-9, atol=1e-9 * float(ref.abs().max()))


def test_noalloc_python_wrapper_and_limit():
    from hipnufft import _capi

    a, b = _pair(5, 4096, 3)
    np.testing.assert_array_equal(_capi.dot_c64_rows_noalloc(a, b).cpu().numpy(),
                                  _capi.dot_c64_rows(a, b).cpu().numpy())
    big_a, big_b = _pair(_capi.NOALLOC_DOT_ROWS + 1, 16, 4)
    out = torch.empty((_capi.NOALLOC_DOT_ROWS + 1, 2), dtype=torch.float64, device=DEV)
    rc = _noalloc(big_a, big_b, True, out)
    assert rc == _capi.ERR_ARG
    msg = _capi.lib().hipnufft_last_error().decode()
    assert "at most" in msg
    with pytest.raises(ValueError, match="at most"):
        _capi.dot_c64_rows_noalloc(big_a, big_b)
