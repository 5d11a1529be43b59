"""Least-squares solvers on the device (``src/mrinufft/extras/optim.py``).

* ``cg``   restates ``optim.py:801-902`` step for step: Lipschitz step size
  computed with the density, density-compensated warm start
  (``_scaled_dcp``, :151-165) then density removed during the iterations,
  ``velocity = tol * velocity + grad / L`` first step, Polak-Ribiere beta
  with the *non-conjugated* dot of ``numpy.dot`` (:875-877) and the
  lexicographic complex ``max(0, beta)`` (:878), stop when
  ``||grad|| <= tol`` (:872).  The two dots and the norm are accumulated in
  FP64 by ``hipnufft_dot_c64``.  With ``graph=True`` one iteration is
  captured in a HIP graph and the stopping test becomes a device flag that
  predicates the update (the result is identical to breaking out).
* ``lsqr`` / ``lsmr`` restate :249-495 / :498-798 (per-batch scalars kept on
  the host, vectors on the device).
"""

from __future__ import annotations

import inspect
import math

import numpy as np
import torch

from . import _capi
from ._array import from_device, kind_of, to_device
from ._registry import MethodRegister

register_optim = MethodRegister("optimizer")
get_optimizer = register_optim.make_getter()


def _dev(op):
    return getattr(op, "device", torch.device("cuda"))


def _dot(a, b, conj):
    return _capi.dot_c64(a.reshape(-1), b.reshape(-1), conj=conj)


def _lexi_pos(re, im):
    """``complex(re, im) > 0`` with numpy's lexicographic complex ordering."""
    return re > 0 or (re == 0 and im > 0)


def _scaled_dcp(operator, kspace_data):
    """Density-compensated adjoint scaled to the data norm (``optim.py:151-165``)."""
    x_init = operator.adj_op(kspace_data)
    y = operator.op(x_init)
    return x_init * (torch.linalg.vector_norm(kspace_data) / torch.linalg.vector_norm(y))


def _with_damp(grad, image, damp, x0, shape):
    if damp:
        if x0 is not None:
            grad = grad + damp * (image - x0.reshape(shape))
        else:
            grad = grad + damp * image
    return grad


@register_optim
def cg(operator, kspace_data, damp=0.0, x0=None, x_init=None, max_iter=10, tol=1e-4,
       progressbar=False, callback=None, graph=False, toeplitz=False):
    """Non-linear CG of ``optim.py:801-902`` on the device (FP64 inner products).

    ``toeplitz=True`` evaluates each gradient ``A^H D (A x - y)`` as
    ``G x - A^H D y`` with the operator's Toeplitz Gram form ``G`` (built
    once, ``gram_op``) and ``A^H D y`` computed once: the iterations then run
    FFTs only, no spreading or interpolation, and so do the power iterations
    of the Lipschitz estimate where the operator offers that form.  Same
    recurrences otherwise.
    """
    kind = kind_of(kspace_data)
    dev = _dev(operator)
    y = to_device(kspace_data, torch.complex64, dev)
    if x0 is not None:
        x0 = to_device(x0, torch.complex64, dev)
    if toeplitz and "toeplitz" in inspect.signature(operator.get_lipschitz_cst).parameters:
        lipschitz_cst = float(operator.get_lipschitz_cst(toeplitz=True))
    else:
        lipschitz_cst = float(operator.get_lipschitz_cst())
    old_density = None
    shape = operator.img_full_shape
    sq = operator.squeeze_dims
    operator.squeeze_dims = False
    try:
        if operator.uses_density:
            if x_init is None:
                x_init = _scaled_dcp(operator, y)
            old_density = operator.density
            operator.density = None
        image = (torch.zeros(shape, dtype=torch.complex64, device=dev) if x_init is None
                 else to_device(x_init, torch.complex64, dev).reshape(shape).clone())
        if toeplitz:
            rhs = to_device(operator.adj_op(y), torch.complex64, dev).reshape(shape)

            def dc(img):
                return to_device(operator.gram_op(img), torch.complex64, dev).reshape(shape) - rhs
        else:
            def dc(img):
                return operator.data_consistency(img, y).reshape(shape)
        velocity = torch.zeros_like(image)
        grad = dc(image)
        grad = _with_damp(grad, image, damp, x0, shape)
        velocity = tol * velocity + grad / lipschitz_cst
        image = image - velocity
        if graph and callback is None:
            image = _cg_graph(dc, image, grad, velocity, lipschitz_cst, damp, x0,
                              max_iter, tol, shape)
            callbacks_results = []
        else:
            callbacks_results = []
            for _ in range(max_iter):
                grad_new = dc(image)
                grad_new = _with_damp(grad_new, image, damp, x0, shape)
                nrm2 = _dot(grad_new, grad_new, True)
                num = _dot(grad_new, grad_new - grad, False)
                den = _dot(grad, grad, False)
                vals = torch.cat([nrm2, num, den]).cpu().numpy()  # one host sync
                if math.sqrt(max(vals[0], 0.0)) <= tol:
                    break
                beta = complex(vals[2], vals[3]) / complex(vals[4], vals[5])
                if not _lexi_pos(beta.real, beta.imag):
                    beta = 0
                velocity = grad_new + beta * velocity
                image = image - velocity / lipschitz_cst
                grad = grad_new
                if callback:
                    callbacks_results.append(callback(image, operator, y, damp=damp, x0=x0))
    finally:
        operator.squeeze_dims = sq
        if old_density is not None:
            operator.density = old_density
    if operator.squeeze_dims:
        image = operator._safe_squeeze(image)
    image = from_device(image, kind)
    if callbacks_results:
        return image, callbacks_results
    return image


def _cg_graph(dc, image, grad, velocity, L, damp, x0, max_iter, tol, shape):
    """CG iterations with a device stop flag, one iteration captured in a HIP graph."""
    dev = image.device
    st = dict(image=image.contiguous(), grad=grad.contiguous(), vel=velocity.contiguous(),
              done=torch.zeros((), dtype=torch.bool, device=dev))
    bufs = dict(nrm=torch.empty(2, dtype=torch.float64, device=dev),
                num=torch.empty(2, dtype=torch.float64, device=dev),
                den=torch.empty(2, dtype=torch.float64, device=dev),
                ws=torch.empty(_capi.DOT_WS, dtype=torch.float64, device=dev),
                zero=torch.zeros((), dtype=torch.float64, device=dev))
    inv_l = 1.0 / L
    tol2 = tol * tol

    def step():
        gn = _with_damp(dc(st["image"]), st["image"], damp, x0, shape)
        ws = bufs["ws"]
        _capi.dot_c64(gn.reshape(-1), gn.reshape(-1), True, out=bufs["nrm"], ws=ws)
        diff = gn - st["grad"]
        _capi.dot_c64(gn.reshape(-1), diff.reshape(-1), False, out=bufs["num"], ws=ws)
        _capi.dot_c64(st["grad"].reshape(-1), st["grad"].reshape(-1), False, out=bufs["den"],
                      ws=ws)
        stop = st["done"] | (bufs["nrm"][0] <= tol2)
        n_re, n_im = bufs["num"][0], bufs["num"][1]
        d_re, d_im = bufs["den"][0], bufs["den"][1]
        dd = d_re * d_re + d_im * d_im
        b_re = (n_re * d_re + n_im * d_im) / dd
        b_im = (n_im * d_re - n_re * d_im) / dd
        pos = (b_re > 0) | ((b_re == 0) & (b_im > 0))
        b_re = torch.where(pos, b_re, bufs["zero"])
        b_im = torch.where(pos, b_im, bufs["zero"])
        beta = torch.complex(b_re, b_im).to(torch.complex64)
        vel_new = gn + beta * st["vel"]
        img_new = st["image"] - vel_new * inv_l
        st["image"].copy_(torch.where(stop, st["image"], img_new))
        st["vel"].copy_(torch.where(stop, st["vel"], vel_new))
        st["grad"].copy_(torch.where(stop, st["grad"], gn))
        st["done"].copy_(stop)

    if max_iter <= 0:
        return st["image"]
    # warm-up iteration eagerly (also allocates every workspace), then capture
    step()
    remaining = max_iter - 1
    if remaining > 0:
        s = torch.cuda.Stream(device=dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            # one more eager step on the side stream before capture so lazily
            # created library state is bound to the capture stream
            step()
            remaining -= 1
            torch.cuda.current_stream(dev).synchronize()
            if remaining > 0:
                with torch.cuda.graph(g, stream=s):
                    step()
        torch.cuda.current_stream(dev).wait_stream(s)
        for _ in range(remaining):
            g.replay()
        torch.cuda.current_stream(dev).wait_stream(s)
    return st["image"]


def _norm_b(x):
    return torch.linalg.vector_norm(x.reshape(x.shape[0], -1), dim=1).double().cpu().numpy()


def _bc(v, x):
    t = torch.as_tensor(np.asarray(v, dtype=np.float64), device=x.device).to(torch.float32)
    return t.reshape((-1,) + (1,) * (x.ndim - 1))


def _sym_ortho(a, b):
    """Stable Givens rotation (``optim.py:213-246``) on per-batch numpy arrays."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if np.any(b == 0):
        return np.sign(a), 0, np.abs(a)
    elif np.any(a == 0):
        return 0, np.sign(b), np.abs(b)
    elif np.any(np.abs(b) > np.abs(a)):
        tau = a / b
        s = np.sign(b) / np.sqrt(1 + tau * tau)
        c = s * tau
        r = b / s
    else:
        tau = b / a
        c = np.sign(a) / np.sqrt(1 + tau * tau)
        s = c * tau
        r = a / c
    return c, s, r


def _setup_ls(operator, kspace_data, x0, x_init):
    kind = kind_of(kspace_data)
    dev = _dev(operator)
    y = to_device(kspace_data, torch.complex64, dev)
    old_density = None
    if operator.uses_density:
        if x_init is None:
            sq = operator.squeeze_dims
            operator.squeeze_dims = False
            x_init = _scaled_dcp(operator, y)
            operator.squeeze_dims = sq
        old_density = operator.density
        operator.density = None
    y = y.reshape(operator.ksp_full_shape)
    if x0 is not None:
        x0 = to_device(x0, torch.complex64, dev).reshape(operator.img_full_shape)
    if x_init is None:
        x = (torch.zeros(operator.img_full_shape, dtype=torch.complex64, device=dev)
             if x0 is None else x0.clone())
    else:
        x = to_device(x_init, torch.complex64, dev).reshape(operator.img_full_shape).clone()
    return kind, y, x, x0, old_density


@register_optim
def lsqr(operator, kspace_data, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8, max_iter=100,
         x0=None, x_init=None, callback=None, progressbar=False):
    """LSQR (Paige & Saunders) restated from ``optim.py:249-495``."""
    sq = operator.squeeze_dims
    kind, b, x, x0, old_density = _setup_ls(operator, kspace_data, x0, x_init)
    operator.squeeze_dims = False
    A = lambda v: operator.op(v).reshape(operator.ksp_full_shape)  # noqa: E731
    AT = lambda u: operator.adj_op(u).reshape(operator.img_full_shape)  # noqa: E731
    try:
        ctol = 1 / conlim if conlim > 0 else 0
        eps = np.finfo(np.float32).eps
        u = b.clone()
        bnorm = _norm_b(u)
        beta = bnorm.copy()
        if x0 is not None:
            u = u - A(x0)
            beta = _norm_b(u)
        if np.all(beta > 0):
            u = u / _bc(beta, u)
            v = AT(u)
            alpha = _norm_b(v)
        else:
            v = x.clone()
            alpha = np.zeros(v.shape[0])
        if np.any(alpha * beta == 0):
            return from_device(x, kind)
        if np.all(alpha > 0):
            v = v / _bc(alpha, v)
        w = v.clone()
        rhobar = alpha
        phibar = beta
        ddnorm = res2 = xxnorm = z = anorm = 0.0
        dampsq = damp ** 2
        cs2, sn2 = -1, 0.0
        callback_returns = []
        for _ in range(max_iter):
            u = A(v) - _bc(alpha, u) * u
            beta = _norm_b(u)
            if np.all(beta > 0):
                u = u / _bc(beta, u)
                anorm = np.sqrt(anorm ** 2 + alpha ** 2 + beta ** 2 + dampsq)
                v = AT(u) - _bc(beta, v) * v
                alpha = _norm_b(v)
                if np.all(alpha > 0):
                    v = v / _bc(alpha, v)
            if damp:
                rhobar1 = np.sqrt(rhobar ** 2 + dampsq)
                cs1 = rhobar / rhobar1
                sn1 = damp / rhobar1
                psi = sn1 * phibar
                phibar = cs1 * phibar
            else:
                rhobar1 = rhobar
                psi = 0.0
            cs, sn, rho = _sym_ortho(rhobar1, beta)
            theta = sn * alpha
            rhobar = -cs * alpha
            phi = cs * phibar
            phibar = sn * phibar
            tau = sn * phi
            t1 = phi / rho
            t2 = -theta / rho
            dk = w / _bc(rho, w)
            x = x + _bc(t1, w) * w
            w = w * _bc(t2, w) + v
            ddnorm = ddnorm + _norm_b(dk) ** 2
            delta = sn2 * rho
            gambar = -cs2 * rho
            rhs = phi - delta * z
            zbar = rhs / gambar
            xnorm = np.sqrt(xxnorm + zbar ** 2)
            gamma = np.sqrt(gambar ** 2 + theta ** 2)
            cs2 = gambar / gamma
            sn2 = theta / gamma
            z = rhs / gamma
            xxnorm = xxnorm + z ** 2
            acond = anorm * np.sqrt(ddnorm)
            res1 = phibar ** 2
            res2 = res2 + psi ** 2
            rnorm = np.sqrt(res1 + res2)
            arnorm = alpha * np.abs(tau)
            test1 = rnorm / bnorm
            test2 = arnorm / (anorm * rnorm + eps)
            test3 = 1 / (acond + eps)
            t1 = test1 / (1 + anorm * xnorm / bnorm)
            rtol = btol + atol * anorm * xnorm / bnorm
            if callback:
                callback_returns.append(callback(x, operator, b, damp=damp, x0=x0))
            if (np.all(1 + test3 <= 1) or np.all(1 + test2 <= 1) or np.all(1 + t1 <= 1)
                    or np.all(test3 <= ctol) or np.all(test2 <= atol) or np.all(test1 <= rtol)):
                break
    finally:
        operator.squeeze_dims = sq
        if old_density is not None:
            operator.density = old_density
    if operator.squeeze_dims:
        x = operator._safe_squeeze(x)
    x = from_device(x, kind)
    if callback_returns:
        return x, callback_returns
    return x


@register_optim
def lsmr(operator, kspace_data, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8, max_iter=100,
         x0=None, x_init=None, callback=None, progressbar=False):
    """LSMR (Fong & Saunders) restated from ``optim.py:498-798``."""
    sq = operator.squeeze_dims
    kind, b, x, x0, old_density = _setup_ls(operator, kspace_data, x0, x_init)
    operator.squeeze_dims = False
    A = lambda v: operator.op(v).reshape(operator.ksp_full_shape)  # noqa: E731
    AT = lambda u: operator.adj_op(u).reshape(operator.img_full_shape)  # noqa: E731
    callback_returns = []
    try:
        ctol = 1 / conlim if conlim > 0 else 0
        u = b.clone()
        normb = _norm_b(u)
        beta = normb.copy()
        if x0 is not None:
            u = u - A(x)
            beta = _norm_b(u)
        if np.all(beta > 0):
            u = u / _bc(beta, u)
            v = AT(u)
            alpha = _norm_b(v)
        else:
            v = x.clone()
            alpha = np.zeros(v.shape[0])
        if np.any(alpha * beta == 0):
            return from_device(x, kind)
        if np.all(alpha > 0):
            v = v / _bc(alpha, v)
        dampv = np.full(operator.n_batchs, damp, np.float64)
        itn = 0
        zetabar = alpha * beta
        alphabar = alpha
        rho = rhobar = cbar = 1
        sbar = 0
        h = v.clone()
        hbar = torch.zeros_like(v)
        betadd = beta
        betad = 0
        rhodold = 1
        tautildeold = thetatilde = zeta = d = 0
        normA2 = alpha * alpha
        maxrbar = 0
        minrbar = 1e100
        for _ in range(max_iter):
            u = A(v) - _bc(alpha, u) * u
            beta = _norm_b(u)
            if np.all(beta > 0):
                u = u / _bc(beta, u)
                v = AT(u) - _bc(beta, v) * v
                alpha = _norm_b(v)
                if np.all(alpha > 0):
                    v = v / _bc(alpha, v)
            chat, shat, alphahat = _sym_ortho(alphabar, dampv)
            rhoold = rho
            c, s, rho = _sym_ortho(alphahat, beta)
            thetanew = s * alpha
            alphabar = c * alpha
            rhobarold = rhobar
            zetaold = zeta
            thetabar = sbar * rho
            rhotemp = cbar * rho
            cbar, sbar, rhobar = _sym_ortho(cbar * rho, thetanew)
            zeta = cbar * zetabar
            zetabar = -sbar * zetabar
            hbar = hbar * _bc(-(thetabar * rho / (rhoold * rhobarold)), hbar) + h
            x = x + _bc(zeta / (rho * rhobar), hbar) * hbar
            h = h * _bc(-(thetanew / rho), h) + v
            betaacute = chat * betadd
            betacheck = -shat * betadd
            betahat = c * betaacute
            betadd = -s * betaacute
            thetatildeold = thetatilde
            ctildeold, stildeold, rhotildeold = _sym_ortho(rhodold, thetabar)
            thetatilde = stildeold * rhobar
            rhodold = ctildeold * rhobar
            betad = -stildeold * betad + ctildeold * betahat
            tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold
            taud = (zeta - thetatilde * tautildeold) / rhodold
            d = d + betacheck * betacheck
            normr = np.sqrt(d + (betad - taud) ** 2 + betadd * betadd)
            normA2 = normA2 + beta * beta
            normA = np.sqrt(normA2)
            normA2 = normA2 + alpha * alpha
            maxrbar = np.max(np.maximum(maxrbar, rhobarold))
            if itn > 1:
                minrbar = np.min(np.minimum(minrbar, rhobarold))
            condA = np.mean(np.maximum(maxrbar, rhotemp) / np.minimum(minrbar, rhotemp))
            normar = np.abs(zetabar)
            normx = _norm_b(x)
            test1 = normr / normb
            test2 = normar / (normA * normr) if np.all(normA * normr != 0) else np.inf
            test3 = 1 / condA
            t1 = test1 / (1 + normA * normx / normb)
            rtol = btol + atol * normA * normx / normb
            if callback:
                callback_returns.append(callback(x, operator, b, damp=damp, x0=x0))
            if (np.all(1 + test3 <= 1) or np.all(1 + test2 <= 1) or np.all(1 + t1 <= 1)
                    or np.all(test3 <= ctol) or np.all(test2 <= atol) or np.all(test1 <= rtol)):
                break
    finally:
        operator.squeeze_dims = sq
        if old_density is not None:
            operator.density = old_density
    if operator.squeeze_dims:
        x = operator._safe_squeeze(x)
    x = from_device(x, kind)
    if callback_returns:
        return x, callback_returns
    return x


def loss_l2_reg(image, operator, kspace_data, damp=0.0, x0=None):
    """``||A x - y||`` (+ ``damp^2 ||x - x0||``) per batch (``optim.py:89-148``)."""
    r = operator.op(image).reshape(operator.ksp_full_shape) - kspace_data.reshape(
        operator.ksp_full_shape)
    res = torch.linalg.vector_norm(r.reshape(operator.n_batchs, -1), dim=1).squeeze()
    if damp:
        xi = image.reshape(operator.img_full_shape)
        if x0 is not None:
            xi = xi - x0.reshape(operator.img_full_shape)
        res = res + damp ** 2 * torch.linalg.vector_norm(xi.reshape(operator.n_batchs, -1), dim=1)
    return res


def loss_l2_AHreg(image, operator, kspace_data, *args, **kwargs):
    """``||A^H (A x - y)||`` per batch (``optim.py:168-210``)."""
    r = operator.op(image).reshape(operator.ksp_full_shape) - kspace_data.reshape(
        operator.ksp_full_shape)
    g = operator.adj_op(r).reshape(operator.img_full_shape)
    return torch.linalg.vector_norm(g.reshape(operator.n_batchs, -1), dim=1).squeeze()
