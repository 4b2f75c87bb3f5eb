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


def extinction_with_drude_gradient(bem, enei_list, planewave, chunk=16, inverse=False):
    """Extinction (n_lambda, npol) and its derivatives (n_lambda, npol, 3)
    with respect to (eps0, wp, gammad) of the inside Drude material, for a
    BEMRet solver ``bem`` (double or mixed precision).  With ``inverse`` the
    factors are replaced by explicit inverses (BEMRet.invert_factors) before
    the two solves, which then run as matrix-vector products; off by default:
    two forward solves (about six applications of the windowed triangular
    solves) do not repay the three inversions (2880 faces, 21 wavelengths:
    0.177 against 0.158 s)."""
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
        if inverse:
            bem.invert_factors(F)
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
        # factors held by the solver belong to the old material
        bem._fac, bem.enei, bem._key = None, None, None
        ext, grad = extinction_with_drude_gradient(bem, enei_list, planewave)
        ctx.save_for_backward(torch.as_tensor(grad, dtype=params.dtype, device=params.device))
        return torch.as_tensor(ext, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return (grad_out[:, :, None] * grad).sum(dim=(0, 1)), None, None, None


# ---------------------------------------------------------------------------
# Vertex-position gradient
# ---------------------------------------------------------------------------

_GR_SLOTS = 31  # per-face gradient layout of the kernels: pos 3 | nvec 3 | tvec1 3 | tvec2 3 | area 1 | nodes 18


def _face_geometry_torch(X, flat):
    """pos, nvec, tvec1, tvec2, area of 6-node faces X (n, 6, 3) as torch
    functions of X, the same formulas as Particle._geometry (particle.py
    _norm / _norm_flat; area by the 28-point rule of the surface Jacobian)."""
    from .quadrature import rule18, tri6_tables
    S, Dx, Dy, w = (torch.as_tensor(a, dtype=X.dtype, device=X.device) for a in tri6_tables(*rule18()))
    tx = torch.einsum("qc,ncd->nqd", Dx, X)
    ty = torch.einsum("qc,ncd->nqd", Dy, X)
    jac = torch.linalg.vector_norm(torch.cross(tx, ty, dim=2), dim=2)
    area = (0.5 * w[None, :] * jac).sum(dim=1)
    S0, Dx0, Dy0, _ = (torch.as_tensor(a, dtype=X.dtype, device=X.device)
                       for a in tri6_tables(np.array([1 / 3]), np.array([1 / 3]), np.array([1.0])))
    pos = torch.einsum("c,ncd->nd", S0[0], X)
    a = torch.einsum("c,ncd->nd", Dy0[0], X)
    b = torch.einsum("c,ncd->nd", Dx0[0], X)
    nv = torch.cross(a, b, dim=1)
    nv = nv / torch.linalg.vector_norm(nv, dim=1, keepdim=True)
    if flat:
        t1 = X[:, 0, :] - X[:, 1, :]
    else:
        t1 = a
    t1 = t1 / torch.linalg.vector_norm(t1, dim=1, keepdim=True)
    t2 = torch.cross(nv, t1, dim=1)
    return pos, nv, t1, t2, area


def _chain_to_vertices(cp, Gr):
    """Per-face gradients Gr (L, P, n, 31) -> per particle the gradient with
    respect to its nodes verts2 (L, P, nverts2, 3) and, for a particle built
    from corner triangles (edge midpoints derived), with respect to its
    corners (L, P, nverts, 3)."""
    L, P, n, _ = Gr.shape
    out = []
    G = torch.as_tensor(Gr, dtype=torch.float64).reshape(L * P, n, _GR_SLOTS)
    for ip, part in enumerate(cp.p):
        lo, hi = int(cp.offsets[ip]), int(cp.offsets[ip + 1])
        g = G[:, lo:hi]
        # per-face Jacobian of the 13 geometric outputs (pos, nvec, tvec1,
        # tvec2, area) with respect to the 18 node coordinates, one backward
        # pass per output component (faces are independent), then one
        # contraction with the cotangents of every wavelength and polarisation;
        # grad mode on: this also runs inside autograd.Function.forward
        with torch.enable_grad():
            X = torch.as_tensor(part.nodes, dtype=torch.float64).requires_grad_(True)
            pos, nv, t1, t2, area = _face_geometry_torch(X, part.interp == "flat")
            comps = [pos[:, k] for k in range(3)] + [nv[:, k] for k in range(3)] + \
                    [t1[:, k] for k in range(3)] + [t2[:, k] for k in range(3)] + [area]
            J = torch.stack([torch.autograd.grad(cpt.sum(), X, retain_graph=(k < 12))[0]
                             for k, cpt in enumerate(comps)])                # (13, nf, 6, 3)
        gX = torch.einsum("kfcd,bfk->bfcd", J, g[:, :, 0:13]) + g[:, :, 13:31].reshape(L * P, hi - lo, 6, 3)
        nv2 = len(part.verts2)
        gv2 = torch.zeros((L * P, nv2, 3), dtype=torch.float64)
        idx = torch.as_tensor(part.faces6.reshape(-1), dtype=torch.int64)
        gv2.index_add_(1, idx, gX.reshape(L * P, -1, 3))
        gv2 = gv2.reshape(L, P, nv2, 3).numpy()
        res = {"nodes": gv2}
        if part.mid_edges is not None:
            nv = len(part.verts)
            gc = gv2[:, :, :nv].copy()
            gm = 0.5 * gv2[:, :, nv:]
            np.add.at(gc, (slice(None), slice(None), part.mid_edges[:, 0]), gm)
            np.add.at(gc, (slice(None), slice(None), part.mid_edges[:, 1]), gm)
            res["verts"] = gc
        out.append(res)
    return out


def extinction_with_shape_gradient(bem, enei_list, planewave, chunk=16, per_face=False, inverse=True):
    """Extinction (n_lambda, npol) and its derivatives with respect to the
    mesh: one adjoint solve A^T y = c per wavelength with the factors of the
    forward solve (c the optical-theorem functional), then the HIP kernels
    contract y and x with the derivatives of the Green entries (far-field
    pairs, refined near-field and self pairs, the closed-surface correction),
    of the plane-wave right side and of c, per face geometric input; the
    chain rule to the nodes runs in torch.  Returns (ext, grads) with grads a
    list per particle of dicts: 'nodes' (n_lambda, npol, nverts2, 3) and, for
    particles built from corner triangles, 'verts' (n_lambda, npol, nverts, 3).
    With ``per_face`` the per-face gradients (n_lambda, npol, n, 31) are
    returned as a third value.  With ``inverse`` the factors are replaced by
    explicit inverses (BEMRet.invert_factors) before the forward and adjoint
    solves (each applies them about three times), which then run as
    matrix-vector products instead of triangular solves: the adjoint's
    row-major triangular steps (~10 ms per solve at 2880 faces and 21
    wavelengths) cost more than the three inversions (~29 ms)."""
    if not isinstance(bem, BEMRet):
        raise TypeError("extinction_with_shape_gradient: expects a BEMRet solver")
    p, g = bem.p, bem.g
    if getattr(g, "has_quads", False):
        raise NotImplementedError("extinction_with_shape_gradient: surfaces of 6-node triangles only "
                                  "(this particle has quadrilateral faces)")
    enei_list = np.atleast_1d(np.asarray(enei_list, dtype=float))
    P = planewave.npol
    if P > 4:
        raise ValueError("extinction_with_shape_gradient: at most 4 polarisations")
    dev = g.device
    st = nat.stream_handle(dev)
    n = g.n
    inm, outm, pol, dr = planewave._side_flags(p, g)
    exts, grs = [], []
    for c0 in range(0, len(enei_list), chunk):
        lam = enei_list[c0:c0 + chunk]
        B = len(lam)
        F = bem.factor(lam)
        if inverse:
            bem.invert_factors(F)
        R = planewave.rhs(p, lam, g)
        X = bem.solve_batch(F, R)
        ext = planewave.extinction_batch(p, lam, X, g)
        pw = planewave._pw(p, lam, dev)
        C = torch.empty_like(X)
        nat.call("mnpbem_ext_functional", n, P, nat.ptr(g.pos), nat.ptr(g.area), nat.ptr(inm), nat.ptr(outm),
                 nat.ptr(pol), nat.ptr(dr), nat.ptr(pw), nat.ptr(C), 8 * P, n * 8 * P, B, st)
        Y = bem.solve_adjoint_batch(F, C)
        G64, _ = bem.green64(F)
        Q = torch.empty_like(X)
        for s in range(2):
            nat.call("mnpbem_gemv_t_batched", nat.DTYPE_C128, n, 4 * P, nat.ptr(G64[s]), n, n * n,
                     nat.ptr_off(X, s * 4 * P), 8 * P, n * 8 * P, nat.ptr_off(Q, s * 4 * P), 8 * P, n * 8 * P,
                     1.0, 0.0, 0.0, B, st)
        del G64
        Lw = torch.empty((B, n, 2, P, 8), dtype=torch.complex128, device=dev)
        Rw = torch.empty((B, n, 2, P, 4), dtype=torch.complex128, device=dev)
        Wc = torch.empty((B, n, P), dtype=torch.complex128, device=dev)
        Gr = torch.zeros((B, P, n, _GR_SLOTS), dtype=torch.float64, device=dev)
        nat.call("mnpbem_grad_prep", n, P, nat.ptr(X), nat.ptr(Y), 8 * P, n * 8 * P, nat.ptr(g.nvec), nat.ptr(F.par),
                 nat.ptr(Lw), nat.ptr(Rw), nat.ptr(Wc), B, st)
        nat.call("mnpbem_grad_far", n, P, nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(g.area), nat.ptr(g.mat[0]),
                 nat.ptr(g.mat[1]), nat.ptr(g._surf), nat.ptr(F.kmat), g.nmat, nat.ptr(g._rowptr()),
                 nat.ptr(g.pcol), nat.ptr(g._colptr), nat.ptr(g._crow), nat.ptr(Lw), nat.ptr(Rw), nat.ptr(Wc),
                 nat.ptr(Gr), B, st)
        Ws = torch.empty((B * P * 11 * g.npairs * 5,), dtype=torch.float64, device=dev)
        nat.call("mnpbem_grad_refined", g.order, n, P, g.npairs, nat.ptr(g.prow), nat.ptr(g.pcol),
                 nat.ptr(g._pkind), nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(g._tv1), nat.ptr(g._tv2),
                 nat.ptr(g.area), nat.ptr(g._nodes), nat.ptr(g.mat[0]), nat.ptr(g.mat[1]), nat.ptr(g._surf),
                 nat.ptr(F.kmat), g.nmat, nat.ptr(g._qn), g._qn.shape[0], nat.ptr(g._qs), g._qs.shape[0],
                 nat.ptr(Lw), nat.ptr(Rw), nat.ptr(Wc), nat.ptr(g._rowptr()), nat.ptr(g._colptr),
                 nat.ptr(g._cpair), nat.ptr(Ws), Ws.numel(), nat.ptr(Gr), B, st)
        del Ws
        nat.call("mnpbem_grad_local", n, P, nat.ptr(X), nat.ptr(Y), nat.ptr(Q), nat.ptr(R), 8 * P, n * 8 * P,
                 nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(g.area), nat.ptr(inm), nat.ptr(outm), nat.ptr(pol),
                 nat.ptr(dr), nat.ptr(pw), nat.ptr(g.corr), nat.ptr(Wc), nat.ptr(Gr), B, st)
        exts.append(ext.cpu().numpy())
        grs.append(Gr.cpu().numpy())
        del F, X, Y, C, Q, Lw, Rw, Wc, Gr, R
    ext = np.concatenate(exts)
    Gr = np.concatenate(grs)
    grads = _chain_to_vertices(p, Gr)
    return (ext, grads, Gr) if per_face else (ext, grads)


class ShapeExtinction(torch.autograd.Function):
    """Extinction spectrum as a differentiable function of the vertex
    positions of a one-particle ComParticle: ``ext = ShapeExtinction.apply(
    verts, build, enei_list, planewave, precision)`` with ``verts`` a float64
    tensor (nverts, 3) and ``build(verts_numpy) -> ComParticle`` rebuilding
    the particle (corner triangles or 6-node faces); ``ext`` is (n_lambda,
    npol).  The backward pass uses the gradient of the adjoint computed with
    the forward solve."""

    @staticmethod
    def forward(ctx, verts, build, enei_list, planewave, precision="double"):
        cp = build(verts.detach().cpu().numpy().astype(float))
        if cp.np != 1:
            raise NotImplementedError("ShapeExtinction: one particle")
        ext, grads = extinction_with_shape_gradient(BEMRet(cp, precision=precision), enei_list, planewave)
        gv = grads[0]["verts"] if grads[0].get("verts") is not None else grads[0]["nodes"]
        if gv.shape[2] != verts.shape[0]:
            raise ValueError("ShapeExtinction: build() must keep the vertex count")
        ctx.save_for_backward(torch.as_tensor(gv, dtype=verts.dtype, device=verts.device))
        return torch.as_tensor(ext, dtype=verts.dtype, device=verts.device)

    @staticmethod
    def backward(ctx, grad_out):
        (gv,) = ctx.saved_tensors
        return (grad_out[:, :, None, None] * gv).sum(dim=(0, 1)), None, None, None, None
