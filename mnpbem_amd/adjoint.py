"""Derivatives of the extinction spectrum with respect to the Drude
parameters of the particle material, on the GPU.

ext = alpha Im(c^T x) with A(eps1) x = b, where eps1 is the permittivity
inside the particle, c the forward far-field functional of the plane wave
and b independent of eps1 (the wave comes from the embedding medium).  x is
holomorphic in eps1, so one extra solve with the factors already computed
for the spectrum,

    A x' = -(dA/deps1) x,      d ext / dp = Im(alpha c^T x' deps1/dp),

gives the derivative for every real parameter p of the material (eps0, wp,
gammad of EpsDrude; the derivative of eps is the closed form of
mnpbem/materials/eps_drude.py:73-107).  dA/deps1 has the explicit eps1 terms
of GdA&H Eqs. (14, 17) and the dependence of G1, H1 on k1 = k0 sqrt(eps1)
(dk1/deps1 = k1 / (2 eps1)), assembled on the GPU as dG1/dk1, dH1/dk1.
``DrudeExtinction`` wraps this as a torch.autograd.Function.
"""

import numpy as np
import torch

from . import _native as nat
from .bem import BEMRet
from .materials import EpsDrude


def _inside_drude(p):
    io = p.inout
    m_in = int(io[0, 0])
    if np.any(io[:, 0] != m_in):
        raise NotImplementedError("Drude gradient: every particle must have the same inside material")
    eps = p.eps[m_in - 1]
    if not isinstance(eps, EpsDrude):
        raise TypeError("Drude gradient: the inside material must be an EpsDrude")
    if np.any(io[:, 1] == m_in):
        raise NotImplementedError("Drude gradient: the Drude material must not be the embedding medium")
    return m_in, eps


def extinction_with_drude_gradient(bem, enei_list, planewave, chunk=16):
    """Extinction (n_lambda, npol) and its derivatives (n_lambda, npol, 3)
    with respect to (eps0, wp, gammad) of the inside Drude material, for a
    BEMRet solver ``bem`` (double or mixed precision)."""
    if not isinstance(bem, BEMRet):
        raise TypeError("extinction_with_drude_gradient: expects a BEMRet solver")
    p, g = bem.p, bem.g
    m_in, drude = _inside_drude(p)
    enei_list = np.atleast_1d(np.asarray(enei_list, dtype=float))
    dev = g.device
    st = nat.stream_handle(dev)
    n = g.n
    exts, grads = [], []
    for c0 in range(0, len(enei_list), chunk):
        lam = enei_list[c0:c0 + chunk]
        B = len(lam)
        F = bem.factor(lam)
        R = planewave.rhs(p, lam, g)
        X = bem.solve_batch(F, R)
        P = R.shape[2] // 8
        ext = planewave.extinction_batch(p, lam, X, g)
        # G1, H1 of the inside side in complex128 (kept by the mixed path; the
        # complex128 path factors its copies in place, so assemble again)
        kmat = g.wavenumbers(lam)
        if getattr(F, "G64", None) is not None:
            G1, H1 = F.G64[0], F.H64[0]
        else:
            G1 = torch.empty((B, n, n), dtype=torch.complex128, device=dev)
            H1 = torch.empty_like(G1)
            g.assemble(0, kmat, G1, H1, 1.0, nat.DTYPE_C128)
        dG = torch.empty((B, n, n), dtype=torch.complex128, device=dev)
        dH = torch.empty_like(dG)
        nat.call("mnpbem_green_dk", n, nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(g.area), nat.ptr(g.mat[0]),
                 nat.ptr(kmat), g.nmat, g.order, g.npairs, nat.ptr(g.prow), nat.ptr(g.pcol), nat.ptr(g.g),
                 nat.ptr(g.f), nat.ptr(dG), nat.ptr(dH), n, n * n, B, st)
        C = 8 * P
        Y = torch.empty((B, n, 16 * P), dtype=torch.complex128, device=dev)
        for q, M in enumerate((G1, H1, dG, dH)):
            nat.call("mnpbem_gemv_t_batched", nat.DTYPE_C128, n, 4 * P, nat.ptr(M), n, n * n, nat.ptr(X), C,
                     n * C, nat.ptr_off(Y, q * 4 * P), 16 * P, n * 16 * P, 1.0, 0.0, 0.0, B, st)
        del dG, dH
        eps1 = np.array([complex(drude(l)[0]) for l in lam])
        k1 = np.array([complex(drude(l)[1]) for l in lam])
        kap = k1 / (2.0 * eps1)
        kap_t = torch.as_tensor(np.stack([kap.real, kap.imag], axis=1).copy(), dtype=torch.float64).to(dev)
        R2 = torch.empty_like(R)
        nat.call("mnpbem_deps_rhs", n, P, nat.ptr(Y), 16 * P, n * 16 * P, nat.ptr(R2), C, n * C,
                 nat.ptr(g.nvec), nat.ptr(F.par), nat.ptr(kap_t), B, st)
        X2 = bem.solve_batch(F, R2)
        w_im = planewave.extinction_batch(p, lam, X2, g).cpu().numpy()
        # ext is real-linear in x: ext(i x') = alpha Re(c^T x')
        iunit = torch.zeros((B, P, 2), dtype=torch.float64, device=dev)
        iunit[:, :, 1] = 1.0
        nat.call("mnpbem_mv_scale_batched", n * C, P, 0, nat.ptr(iunit), nat.ptr(X2), n * C, nat.ptr(X2), n * C, B,
                 st)
        w_re = planewave.extinction_batch(p, lam, X2, g).cpu().numpy()
        w = w_re + 1j * w_im                                   # alpha c^T x' per (lambda, pol)
        de = drude.deriv(lam)                                  # (B, 3)
        grads.append(np.imag(w[:, :, None] * de[:, None, :]))
        exts.append(ext.cpu().numpy())
        del F, X, X2, R2, Y
    return np.concatenate(exts), np.concatenate(grads)


class DrudeExtinction(torch.autograd.Function):
    """Extinction spectrum as a differentiable function of the Drude
    parameters (eps0, wp, gammad) of a particle's inside material:
    ``ext = DrudeExtinction.apply(params, bem, enei_list, planewave)`` with
    ``params`` a float64 tensor of shape (3,); ``ext`` is (n_lambda, npol).
    The backward pass uses the derivatives computed alongside the forward
    solve (one extra solve per wavelength with the same factors)."""

    @staticmethod
    def forward(ctx, params, bem, enei_list, planewave):
        _, drude = _inside_drude(bem.p)
        vals = params.detach().cpu().numpy().astype(float)
        drude.eps0, drude.wp, drude.gammad = float(vals[0]), float(vals[1]), float(vals[2])
        ext, grad = extinction_with_drude_gradient(bem, enei_list, planewave)
        ctx.save_for_backward(torch.as_tensor(grad, dtype=params.dtype, device=params.device))
        return torch.as_tensor(ext, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return (grad_out[:, :, None] * grad).sum(dim=(0, 1)), None, None, None
