"""Composite retarded Green function with refined near-field quadrature.

``CompGreenRet(p, p)`` mirrors mnpbem/greenfun/compgreen_ret.py: at
construction the refinement classification (refine_utils.py:29-158), the
refinement coefficients and the Fuchs-Liu closed-surface correction run on
the GPU (HIP kernels in csrc/green.hip), and all of it stays resident in
device memory.
``eval(i, j, key, enei)`` assembles a full matrix on the GPU.
"""

import math

import numpy as np
import torch

from . import _native as nat
from .quadrature import kernel_table, kernel_table9, polar_rule, polar_rule_quad, rule18


def refinement_pairs(pos, brad, rel_cutoff=3.0, abs_cutoff=0.0, device="cuda"):
    """Face pairs needing refined integration (refine_utils.py:29-158),
    classified on the GPU: kind 2 where the centroid distance is zero, kind 1
    where (d - r_j)/r_i < RelCutoff or d - r_j < AbsCutoff.  Returns device
    int32 tensors row, col, kind sorted by (row, col)."""
    dev = torch.device(device)
    n = len(pos)
    tp = _dev(pos, torch.float64, dev)
    tb = _dev(brad, torch.float64, dev)
    st = nat.stream_handle(dev)
    cnt = torch.empty((n,), dtype=torch.int32, device=dev)
    nat.call("mnpbem_refine_classify", n, nat.ptr(tp), nat.ptr(tb), float(rel_cutoff), float(abs_cutoff), 0,
             nat.ptr(cnt), None, None, None, None, st)
    c = cnt.cpu().numpy().astype(np.int64)
    off = torch.as_tensor(np.concatenate([[0], np.cumsum(c)[:-1]])).to(dev)
    total = int(c.sum())
    row = torch.empty((total,), dtype=torch.int32, device=dev)
    col = torch.empty_like(row)
    kind = torch.empty_like(row)
    nat.call("mnpbem_refine_classify", n, nat.ptr(tp), nat.ptr(tb), float(rel_cutoff), float(abs_cutoff), 1,
             None, nat.ptr(off), nat.ptr(row), nat.ptr(col), nat.ptr(kind), st)
    return row, col, kind


def _dev(a, dtype, device):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device)


class CompGreenRet:
    """Retarded Green function G, F, H1 = F + 2 pi, H2 = F - 2 pi between the
    faces of a composite particle and itself (deriv='cart', refinement order
    ``order``, RelCutoff/AbsCutoff as in bemoptions)."""

    name = "greenfunction"
    needs = {"sim": "ret"}

    def __init__(self, p1, p2, order=5, RelCutoff=3, AbsCutoff=0, device="cuda", **options):
        if p1 is not p2:
            raise NotImplementedError("CompGreenRet: only the BEM case p1 is p2 is supported")
        if options.get("deriv", "cart") != "cart":
            raise NotImplementedError("CompGreenRet: only deriv='cart' is supported")
        self.p1 = self.p2 = p = p1
        self.order = int(order)
        self.device = torch.device(device)
        dev = self.device
        n = p.n
        self.n = n
        self.prow, self.pcol, self._pkind = refinement_pairs(p.pos, p.brad, RelCutoff, AbsCutoff, dev)
        row, col, kind = (t.cpu().numpy().astype(np.int64) for t in (self.prow, self.pcol, self._pkind))
        self.npairs = len(row)
        self.row, self.col, self.kind = row, col, kind
        f64, i32 = torch.float64, torch.int32
        self.pos = _dev(p.pos, f64, dev)
        self.nvec = _dev(p.nvec, f64, dev)
        self.area = _dev(p.area, f64, dev)
        self._tv1 = _dev(p.tvec1, f64, dev)
        self._tv2 = _dev(p.tvec2, f64, dev)
        self._nodes = _dev(p.nodes, f64, dev)
        # nodes in the 9-node layout of mixed surfaces (triangles: corners
        # 0-2, midpoints 4-6), taken from the 6-node array when every face is
        # a triangle
        nvert = getattr(p, "nvert", None)
        self.has_quads = nvert is not None and bool(np.any(nvert == 4))
        if self.has_quads:
            n9 = p.nodes9
        else:
            n9 = np.zeros((n, 9, 3))
            n9[:, [0, 1, 2, 4, 5, 6]] = p.nodes
            nvert = np.full(n, 3, dtype=np.int32)
        self._nodes9 = _dev(n9, f64, dev)
        self._nvert = _dev(nvert, i32, dev)
        # per side: 0-based material index of each face
        self.mat = [_dev(p.inout_faces[:, s] - 1, i32, dev) for s in (0, 1)]
        self.nmat = len(p.eps)
        qn = _dev(kernel_table(*rule18()), f64, dev)
        qs = _dev(kernel_table(*polar_rule((7, 5))), f64, dev)
        qq = _dev(kernel_table9(*polar_rule_quad((7, 5))), f64, dev) if self.has_quads else None
        o1 = self.order + 1
        self.g = torch.empty((self.npairs, o1), dtype=f64, device=dev)
        self.f = torch.empty((self.npairs, 3, o1), dtype=f64, device=dev)
        self.fstat = torch.empty((self.npairs,), dtype=f64, device=dev)
        st = nat.stream_handle(dev)
        nat.call("mnpbem_refine_coeffs", self.order, self.npairs, nat.ptr(self.prow), nat.ptr(self.pcol),
                 nat.ptr(self._pkind), nat.ptr(self.pos), nat.ptr(self.nvec), nat.ptr(self._tv1),
                 nat.ptr(self._tv2), nat.ptr(self.pos), nat.ptr(self._nodes9), nat.ptr(self._nvert), nat.ptr(qn),
                 qn.shape[0], nat.ptr(qs), qs.shape[0], nat.ptr(qq), 0 if qq is None else qq.shape[0],
                 nat.ptr(self.g), nat.ptr(self.f), nat.ptr(self.fstat), st)
        # closed-surface correction (each particle closed, outward normals)
        oc = np.lexsort((row, col))
        crow, cpair = row[oc], oc
        colptr = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n))])
        surf = p.surface
        lo = p.offsets[surf]
        hi = p.offsets[surf + 1]
        self.corr = torch.empty((n,), dtype=f64, device=dev)
        keep = [_dev(a, i32, dev) for a in (lo, hi, colptr, crow, cpair)]
        # kept for the geometry derivatives: refined rows per column, the
        # closed surface of each face and the quadrature tables
        self._colptr, self._crow, self._cpair = keep[2], keep[3], keep[4]
        self._surf = _dev(surf, i32, dev)
        self._qn, self._qs = qn, qs
        nat.call("mnpbem_closed_correction", n, nat.ptr(self.pos), nat.ptr(self.nvec), nat.ptr(self.area),
                 *[nat.ptr(t) for t in keep], nat.ptr(self.fstat), 1.0, self.order, nat.ptr(self.f),
                 nat.ptr(self.corr), st)
        torch.cuda.current_stream(dev).synchronize()
        self.con = self._connect()
        _GREEN_CACHE[id(p)] = self

    def _connect(self):
        """Connectivity of regions (compgreen_ret.py:435-505): con[i][j][a, b]
        = material index when side i of particle a and side j of particle b
        hold the same material, else 0."""
        io = self.p1.inout
        return [[np.where(io[:, i][:, None] == io[:, j][None, :], io[:, i][:, None], 0)
                 for j in range(2)] for i in range(2)]

    def wavenumbers(self, enei_list):
        """(B, nmat) complex wavenumbers as a (B, nmat, 2) float64 device tensor."""
        k = np.array([[complex(e(lam)[1]) for e in self.p1.eps] for lam in enei_list])
        return _dev(np.stack([k.real, k.imag], axis=-1), torch.float64, self.device)

    def assemble(self, side, kmat, G, H, hsign, dtype, G32=None, H32=None):
        """Fill transposed G, H buffers (B, n, n) of one side for a batch of
        wavelengths (kmat from ``wavenumbers``)."""
        B = kmat.shape[0]
        n = self.n
        dual = 1 if G32 is not None else 0
        nat.call("mnpbem_green_assemble", dtype, dual, n, nat.ptr(self.pos), nat.ptr(self.nvec),
                 nat.ptr(self.area), nat.ptr(self.mat[side]), nat.ptr(kmat), self.nmat, self.order,
                 self.npairs, nat.ptr(self.prow), nat.ptr(self.pcol), nat.ptr(self.g), nat.ptr(self.f),
                 float(hsign), nat.ptr(G), nat.ptr(H), n, n * n, nat.ptr(G32), nat.ptr(H32), 0, None, B,
                 nat.stream_handle(self.device))

    _KEYS = {"G": (0, 0.0), "F": (1, 0.0), "H1": (1, 1.0), "H2": (1, -1.0),
             "Gp": (2, 0.0), "H1p": (2, 1.0), "H2p": (2, -1.0)}

    def eval(self, i, j, key, enei, ind=None):
        """Green function of regions (i, j) (0 = inside, 1 = outside) at
        wavelength ``enei`` (compgreen_ret.py:507-553): G, F, H1 = F + 2 pi,
        H2 = F - 2 pi as (n, n) complex128 device tensors; the Cartesian
        derivatives Gp, H1p = Gp + 2 pi n, H2p = Gp - 2 pi n as (n, 3, n); the
        int 0 when no particle connects the two regions.  With ``ind``
        (row-major linear indices i * n + j into the (n, n) matrix, as
        compgreen_ret.py:1927-1966 reads them) only those entries: (len(ind),)
        for G, F, H1, H2 and (len(ind), 3) for Gp, H1p, H2p, the three
        derivative components of each selected entry (MATLAB
        Greenfun/@greenret/private/eval2.m)."""
        if key not in self._KEYS:
            raise ValueError("Unknown key: %s" % key)
        con = self.con[i][j]
        if not np.any(con):
            return 0
        if i != j:
            raise NotImplementedError("CompGreenRet.eval: cross-region blocks (nested particles)")
        kind, hs = self._KEYS[key]
        n = self.n
        kmat = self.wavenumbers([enei])
        st = nat.stream_handle(self.device)
        if ind is not None:
            idx = np.asarray(ind, dtype=np.int64).ravel()
            if idx.size and (idx.min() < 0 or idx.max() >= n * n):
                raise IndexError("CompGreenRet.eval: ind outside the %d x %d matrix" % (n, n))
            rows = torch.as_tensor((idx // n).astype(np.int32)).to(self.device)
            cols = torch.as_tensor((idx % n).astype(np.int32)).to(self.device)
            out = torch.empty((len(idx), 3) if kind == 2 else (len(idx),), dtype=torch.complex128,
                              device=self.device)
            nat.call("mnpbem_green_select", n, kind, hs, nat.ptr(self.pos), nat.ptr(self.nvec), nat.ptr(self.area),
                     nat.ptr(self.mat[i]), nat.ptr(kmat), self.nmat, self.order, nat.ptr(self._rowptr()),
                     nat.ptr(self.pcol), nat.ptr(self.g), nat.ptr(self.f), len(idx), nat.ptr(rows), nat.ptr(cols),
                     nat.ptr(out), st)
            return out
        if kind == 2:
            out = torch.empty((1, n, 3, n), dtype=torch.complex128, device=self.device)
            nat.call("mnpbem_green_deriv", n, hs, nat.ptr(self.pos), nat.ptr(self.nvec), nat.ptr(self.area),
                     nat.ptr(self.mat[i]), nat.ptr(kmat), self.nmat, self.order, nat.ptr(self._rowptr()),
                     nat.ptr(self.pcol), nat.ptr(self.g), nat.ptr(self.f), nat.ptr(out), 1, st)
            return out[0]
        G = torch.empty((1, n, n), dtype=torch.complex128, device=self.device)
        H = torch.empty_like(G)
        self.assemble(i, kmat, G, H, hs, nat.DTYPE_C128)
        M = G if key == "G" else H
        return M[0].transpose(0, 1)

    # ------------------------------------------------------------------
    def _rowptr(self):
        if getattr(self, "_rowptr_t", None) is None:
            rp = np.concatenate([[0], np.cumsum(np.bincount(self.row, minlength=self.n))]).astype(np.int64)
            self._rowptr_t = _dev(rp, torch.int64, self.device)
        return self._rowptr_t

    @staticmethod
    def _sig_x(sig):
        X = sig.get("_X")
        if X is None:
            raise ValueError("CompGreenRet: CompStruct does not come from BEMRet.solve / BEMRetIter.solve")
        return X

    def potential(self, sig, inout=1):
        """Potentials and surface derivatives at the faces, inside (inout=1)
        or outside (inout=2) (compgreen_ret.py:971-1038): phi = G sig,
        phip = H sig, a = G h, ap = H h with the Green functions of that side."""
        if inout not in (1, 2):
            raise ValueError("inout must be 1 or 2")
        X = self._sig_x(sig)
        B, n, C = X.shape
        P = C // 8
        s = inout - 1
        kmat = self.wavenumbers([sig.enei])
        G = torch.empty((1, n, n), dtype=torch.complex128, device=self.device)
        H = torch.empty_like(G)
        self.assemble(s, kmat, G, H, 1.0 if s == 0 else -1.0, nat.DTYPE_C128)
        st = nat.stream_handle(self.device)
        Y = torch.empty((2, n, 4 * P), dtype=torch.complex128, device=self.device)
        for q, M in enumerate((G, H)):
            nat.call("mnpbem_gemv_t_batched", nat.DTYPE_C128, n, 4 * P, nat.ptr(M), n, n * n,
                     nat.ptr_off(X, s * 4 * P), C, n * C, nat.ptr(Y[q]), 4 * P, n * 4 * P, 1.0, 0.0, 0.0, 1, st)
        phi, a = Y[0, :, :P], Y[0, :, P:].reshape(n, 3, P)
        phip, ap = Y[1, :, :P], Y[1, :, P:].reshape(n, 3, P)
        if P == 1:
            phi, phip, a, ap = phi[:, 0], phip[:, 0], a[:, :, 0], ap[:, :, 0]
        from .bem import CompStruct
        t = str(inout)
        return CompStruct(self.p1, sig.enei, **{"phi" + t: phi, "phi" + t + "p": phip, "a" + t: a,
                                                "a" + t + "p": ap})

    def field(self, sig, inout=1):
        """Electric and magnetic fields at the faces, inside (inout=1) or
        outside (inout=2) (compgreen_ret.py:912-969), computed on the GPU
        without forming the Cartesian-derivative matrices."""
        if inout not in (1, 2):
            raise ValueError("inout must be 1 or 2")
        X = self._sig_x(sig)
        B, n, C = X.shape
        P = C // 8
        if P > 4:
            raise ValueError("CompGreenRet.field: at most 4 polarisations per call")
        s = inout - 1
        kmat = self.wavenumbers([sig.enei])
        k0 = torch.tensor([2.0 * np.pi / sig.enei], dtype=torch.float64, device=self.device)
        E = torch.empty((1, n, 3 * P), dtype=torch.complex128, device=self.device)
        Hf = torch.empty_like(E)
        nat.call("mnpbem_surface_field", n, P, s, 1.0 if s == 0 else -1.0, nat.ptr(self.pos), nat.ptr(self.nvec),
                 nat.ptr(self.area), nat.ptr(self.mat[s]), nat.ptr(kmat), self.nmat, nat.ptr(k0), self.order,
                 nat.ptr(self._rowptr()), nat.ptr(self.pcol), nat.ptr(self.g), nat.ptr(self.f), nat.ptr(X), C,
                 n * C, nat.ptr(E), nat.ptr(Hf), 3 * P, n * 3 * P, 1, nat.stream_handle(self.device))
        e = E[0].reshape(n, 3, P)
        h = Hf[0].reshape(n, 3, P)
        if P == 1:
            e, h = e[:, :, 0], h[:, :, 0]
        from .bem import CompStruct
        return CompStruct(self.p1, sig.enei, e=e, h=h)


_GREEN_CACHE = {}


def green_for(p):
    """The most recent CompGreenRet built for particle ``p`` (built on demand),
    so that PlaneWaveRet shares the device geometry of the BEM solver."""
    g = _GREEN_CACHE.get(id(p))
    if g is None or g.p1 is not p:
        g = CompGreenRet(p, p)
    return g
