"""Dense retarded BEM solver on the GPU, batched over wavelengths.

``BEMRet`` keeps the interface of mnpbem/bem/bem_ret.py (``init(enei)``,
``solve(exc) -> (sig, self)``, ``bem \\ exc`` as ``bem / exc``,
``bem(enei)``) and adds ``solve_sweep`` which factors and solves a whole
wavelength sweep in batched launches.

Per wavelength (bem_ret.py:293-525; MATLAB BEM/@bemret/private/initmat.m):
G1, H1 (inside k) and G2, H2 (outside k) are assembled; with the matrices
stored transposed, LU(G^T) and the LU solve of H^T give Sigma^T = (H G^-1)^T;
Delta^T = Sigma1^T - Sigma2^T is factored and inverted; Sigma^T =
eps1 Sigma1^T - eps2 Sigma2^T + k^2 (eps1 - eps2)^2 (Delta^-1)^T o (n n^T) is
factored.  The solve is GdA&H Eqs. (19)-(20) (bem_ret.py:1569-1831).

Precision.  ``precision='double'`` runs everything in complex128.
``precision='mixed'`` assembles G/H once in complex128 and keeps a complex64
copy; the factorisations and the direct solve run in complex64 with the
GEMMs on bf16 matrix cores (each float split into two bf16 terms, three
products per real product), and the solution is refined in FP64 against the
complex128 8N operator of GdA&H Eqs. (10, 11, 14, 17) until
||R - A X|| <= refine_tol ||R|| for every wavelength (at most
``refine_maxit`` corrections).  ``precision='mixed-fp32'`` is the same with
the complex64 GEMMs on FP32 matrix cores.
"""

import numpy as np
import torch

from . import _native as nat
from .greenfun import CompGreenRet


class CompStruct:
    """Container of named fields with attribute and item access
    (mnpbem/greenfun/compgreen_stat.py:1154)."""

    def __init__(self, p, enei, **fields):
        self.p = p
        self.enei = enei
        self._fields = dict(fields)

    def __getattr__(self, name):
        f = self.__dict__.get("_fields", {})
        if name in f:
            return f[name]
        raise AttributeError(name)

    def __getitem__(self, name):
        if name == "enei":
            return self.enei
        return self._fields[name]

    def get(self, name, default=None):
        return self._fields.get(name, default)

    def set(self, **fields):
        self._fields.update(fields)
        return self

    def keys(self):
        return list(self._fields.keys())


class _Factorization:
    """Device buffers and factors of one batch of wavelengths."""


def _par_tensor(eps_in, eps_out, k0, device):
    par = np.stack([np.real(eps_in), np.imag(eps_in), np.real(eps_out), np.imag(eps_out), k0], axis=1)
    return torch.as_tensor(np.ascontiguousarray(par), dtype=torch.float64).to(device)


class BEMRet:
    """Retarded BEM solver (dense, direct) for a ComParticle."""

    name = "bemsolver"
    needs = {"sim": "ret"}

    def __init__(self, p, enei=None, precision="double", green=None, refine_tol=1e-12, refine_maxit=8,
                 residual="auto", **options):
        if precision not in ("double", "mixed", "mixed-fp32"):
            raise ValueError("BEMRet: precision must be 'double', 'mixed' or 'mixed-fp32'")
        if residual not in ("auto", "matrix_free", "stored"):
            raise ValueError("BEMRet: residual must be 'auto', 'matrix_free' or 'stored'")
        self.p = p
        self.precision = precision
        # FP64 residual of the mixed path: complex128 copies kept from the
        # assembly ('stored', faster: 4 N^2 x 16 B per wavelength) or Green
        # entries recomputed on the fly ('matrix_free', no copies); 'auto'
        # stores when the copies take under 40 % of the free device memory
        self.residual = residual
        self.refine_tol = refine_tol
        self.refine_maxit = refine_maxit
        self.g = green if green is not None else CompGreenRet(p, p, **options)
        self.device = self.g.device
        self.enei = None
        self._fac = None
        self._key = None
        self.last_refine_iters = None
        self.last_residual = None
        io = p.inout
        if p.np > 1 and (len(np.unique(io[:, 0])) > 1 or len(np.unique(io[:, 1])) > 1):
            raise NotImplementedError("BEMRet: one material per side (scalar eps1, eps2) is supported")
        if np.any(self.g.con[0][1]) or np.any(self.g.con[1][0]):
            raise NotImplementedError("BEMRet: nested particles (connected inside/outside regions)")
        if enei is not None:
            self.init(enei)

    # ------------------------------------------------------------------
    def _eps_sides(self, enei_list):
        p = self.p
        m_in, m_out = p.inout[0, 0], p.inout[0, 1]
        e_in = np.array([complex(p.eps[m_in - 1](lam)[0]) for lam in enei_list])
        e_out = np.array([complex(p.eps[m_out - 1](lam)[0]) for lam in enei_list])
        return e_in, e_out

    def _store_copies(self, B, n):
        """Whether the mixed path keeps complex128 copies for its residual."""
        if self.residual == "stored":
            return True
        if self.residual == "matrix_free":
            return False
        return 4 * B * n * n * 16 < 0.4 * torch.cuda.mem_get_info(self.device)[0]

    def factor(self, enei_list, kmat=None, par=None, store=None, interchanges=False):
        """Assemble and factor the BEM matrices for a batch of wavelengths.
        ``kmat``, ``par`` (device tensors of ``g.wavenumbers`` and the
        per-wavelength permittivities) and ``store`` may be given so that no
        host transfer or device query happens (graph capture).

        G1, G2 and Delta are factored without row interchanges, recording
        per matrix whether partial pivoting would have made one (a
        multiplier above 1 in |re| + |im|); when none would, the factors are
        those of partial pivoting (LAPACK getrf).  Sigma, whose partial
        pivoting does interchange rows, is factored with interchanges for
        precision='double'; the mixed precisions factor it without them
        unless a multiplier exceeds 1024 in |re| + |im| (its factors are the
        inner solve of the FP64 refinement, which restores the accuracy and
        checks it: a refinement that misses refine_tol factors again with
        interchanges; the factorisation without interchanges has a backward
        error of ~1e-15 on Sigma in FP64 and a growth below 5 on the test
        geometries).  Outside graph capture a recorded case makes the whole
        batch factor again with partial pivoting (``interchanges=True``); a
        captured step leaves the record in ``F.npflag`` for its caller
        (``needs_interchanges``)."""
        enei_list = np.atleast_1d(np.asarray(enei_list, dtype=float))
        g, dev = self.g, self.device
        B, n = len(enei_list), g.n
        mixed = self.precision != "double"
        cdt = torch.complex64 if mixed else torch.complex128
        dt = {"double": nat.DTYPE_C128, "mixed": nat.DTYPE_C64X, "mixed-fp32": nat.DTYPE_C64}[self.precision]
        st = nat.stream_handle(dev)
        F = _Factorization()
        F.enei = enei_list
        F.B, F.n, F.dtype, F.cdt = B, n, dt, cdt
        kmat = g.wavenumbers(enei_list) if kmat is None else kmat
        e_in, e_out = self._eps_sides(enei_list)
        k0 = 2.0 * np.pi / enei_list
        F.par = _par_tensor(e_in, e_out, k0, dev) if par is None else par
        F.e_in, F.e_out, F.k0 = e_in, e_out, k0
        nn = n * n
        # GG = [G1^T ; G2^T], HH = [H1^T ; H2^T] stacked so both sides factor in one batch
        GG = torch.empty((2, B, n, n), dtype=cdt, device=dev)
        HH = torch.empty((2, B, n, n), dtype=cdt, device=dev)
        F.kmat = kmat
        F.G64 = F.H64 = None
        if store is None:
            store = mixed and self._store_copies(B, n)
        if mixed and store:
            F.G64 = torch.empty((2, B, n, n), dtype=torch.complex128, device=dev)
            F.H64 = torch.empty((2, B, n, n), dtype=torch.complex128, device=dev)
            for s, hs in ((0, 1.0), (1, -1.0)):
                g.assemble(s, kmat, F.G64[s], F.H64[s], hs, nat.DTYPE_C128, GG[s], HH[s])
        elif mixed:
            for s, hs in ((0, 1.0), (1, -1.0)):
                g.assemble(s, kmat, GG[s], HH[s], hs, nat.DTYPE_C64)
        else:
            for s, hs in ((0, 1.0), (1, -1.0)):
                g.assemble(s, kmat, GG[s], HH[s], hs, nat.DTYPE_C128)
        piv = torch.empty((4, B, n), dtype=torch.int32, device=dev)  # G1, G2, Delta, Sigma
        info = torch.empty((4, B), dtype=torch.int32, device=dev)
        npflag = torch.zeros((4, B), dtype=torch.int32, device=dev)  # G1, G2, Delta, Sigma

        def lu_gd(A, pv, inf, flg, nb):
            if interchanges:
                nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(A), n, nn, nat.ptr(pv), n, nat.ptr(inf), nb, st)
            else:
                nat.call("mnpbem_getrf_nopiv_batched", dt, n, nat.ptr(A), n, nn, nat.ptr(pv), n, nat.ptr(inf),
                         nat.ptr(flg), nb, st)

        # inverse diagonal blocks of the factors: the leaves of the many-column
        # triangular solves below and the small-right-hand-side solves
        nblk = (n + 63) // 64
        F.sD = nblk * 2 * 64 * 64
        F.DG = torch.empty((2, B, F.sD), dtype=cdt, device=dev)
        F.DS = torch.empty((B, F.sD), dtype=cdt, device=dev)
        # LU(G^T), then Sigma^T = (G^T)^-1 H^T for both sides in one batch of 2B
        lu_gd(GG, piv, info, npflag, 2 * B)
        nat.call("mnpbem_lu_blockinv_batched", dt, n, nat.ptr(GG), n, nn, nat.ptr(F.DG), F.sD, 2 * B, st)
        # 128-row inverse blocks for the leaves of the many-column solves
        # (complex64 on the split products only; other dtypes ignore them)
        s128 = ((n + 127) // 128) * 2 * 128 * 128
        D128 = None
        if dt == nat.DTYPE_C64X:
            D128 = torch.empty((2 * B, s128), dtype=cdt, device=dev)
            nat.call("mnpbem_lu_blockinv128_batched", dt, n, nat.ptr(GG), n, nn, nat.ptr(F.DG), F.sD,
                     nat.ptr(D128), s128, 2 * B, st)
        nat.call("mnpbem_getrs_blockinv_batched", dt, n, n, nat.ptr(GG), n, nn, nat.ptr(piv), n, nat.ptr(D128),
                 s128, 128, nat.ptr(HH), n, nn, 2 * B, st)
        S1, S2 = HH[0], HH[1]
        Dt = torch.empty((B, n, n), dtype=cdt, device=dev)
        gd = getattr(g, "gdiv", 0)
        nat.call("mnpbem_sigma_batched", dt, 0, n, nat.ptr(S1), nat.ptr(S2), nat.ptr(Dt), None, n, nn,
                 nat.ptr(g.nvec), nat.ptr(F.par), gd, B, st)
        lu_gd(Dt, piv[2], info[2], npflag[2], B)
        # (Delta^-1)^T (LAPACK getri's operation count) in a new buffer, or in
        # place of the factors after interchanges; F.DS holds the inverse
        # diagonal blocks of Delta's factors until Sigma's
        nat.call("mnpbem_lu_blockinv_batched", dt, n, nat.ptr(Dt), n, nn, nat.ptr(F.DS), F.sD, B, st)
        if D128 is not None:
            D128 = D128[:B]
            nat.call("mnpbem_lu_blockinv128_batched", dt, n, nat.ptr(Dt), n, nn, nat.ptr(F.DS), F.sD,
                     nat.ptr(D128), s128, B, st)
        Wk = torch.empty((B, n, n), dtype=cdt, device=dev)
        if interchanges:
            Q = torch.empty((B, n), dtype=torch.int32, device=dev)
            nat.call("mnpbem_getri_blockinv_batched", dt, n, nat.ptr(Dt), n, nn, nat.ptr(piv[2]), n, nat.ptr(D128),
                     s128, 128, nat.ptr(Wk), n, nn, nat.ptr(Q), n, B, st)
            Et = Dt
            del Wk, Q
        else:
            # factors without interchanges (P = I): the inverse is built in Wk
            # and stays there (no permutation, no copy back)
            nat.call("mnpbem_getri_nopiv_blockinv_batched", dt, n, nat.ptr(Dt), n, nn, nat.ptr(D128), s128, 128,
                     nat.ptr(Wk), n, nn, B, st)
            Et = Wk
            del Dt
        del D128
        # Sigma^T overwrites Sigma2^T
        nat.call("mnpbem_sigma_batched", dt, 1, n, nat.ptr(S1), nat.ptr(S2), None, nat.ptr(Et), n, nn,
                 nat.ptr(g.nvec), nat.ptr(F.par), gd, B, st)
        if mixed and not interchanges:
            nat.call("mnpbem_getrf_nopiv_batched", dt, n, nat.ptr(S2), n, nn, nat.ptr(piv[3]), n,
                     nat.ptr(info[3]), nat.ptr(npflag[3]), B, st)
        else:
            nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(S2), n, nn, nat.ptr(piv[3]), n, nat.ptr(info[3]), B, st)
        nat.call("mnpbem_lu_blockinv_batched", dt, n, nat.ptr(S2), n, nn, nat.ptr(F.DS), F.sD, B, st)
        F.GG, F.HH, F.Et, F.piv, F.info = GG, HH, Et, piv, info
        F.npflag = None if interchanges else npflag
        if not interchanges and not torch.cuda.is_current_stream_capturing() and needs_interchanges(npflag):
            del F
            return self.factor(enei_list, kmat=kmat, par=par, store=store, interchanges=True)
        return F

    def invert_factors(self, F):
        """Replace the LU factors of G1^T, G2^T and Sigma^T in ``F`` by the
        inverses (getri in place; the transposed buffers of G1^-1, G2^-1,
        Sigma^-1), for a factorisation that is applied many times to a few
        right-hand sides (the block preconditioner inside GMRES, the forward
        and adjoint solves of the gradients): each solve is then a few
        matrix-vector products instead of chains of latency-bound triangular
        steps.  Work space: one n x n matrix per wavelength while it runs.
        ``solve_batch`` and ``solve_adjoint_batch`` take the inverses from
        here on."""
        if getattr(F, "inverted", False):
            return F
        n, B, dt = F.n, F.B, F.dtype
        nn = n * n
        st = nat.stream_handle(self.device)
        W = torch.empty((B, n, n), dtype=F.cdt, device=self.device)
        Q = torch.empty((B, n), dtype=torch.int32, device=self.device)
        for A, piv, D in ((F.GG[0], F.piv[0], F.DG[0]), (F.GG[1], F.piv[1], F.DG[1]), (F.HH[1], F.piv[3], F.DS)):
            nat.call("mnpbem_getri_blockinv_batched", dt, n, nat.ptr(A), n, nn, nat.ptr(piv), n, nat.ptr(D), F.sD, 64,
                     nat.ptr(W), n, nn, nat.ptr(Q), n, B, st)
        del W, Q
        F.inverted = True
        return F

    def _solve_direct(self, F, R, X, W, P):
        n = F.n
        nn = n * n
        if getattr(F, "inverted", False):
            nat.call("mnpbem_bem_solve_inv_batched", F.dtype, n, P, nat.ptr(F.GG[0]), nat.ptr(F.GG[1]),
                     nat.ptr(F.HH[0]), nat.ptr(F.Et), nat.ptr(F.HH[1]), n, nn, nat.ptr(R), nat.ptr(W), nat.ptr(X),
                     8 * P, n * 8 * P, nat.ptr(self.g.nvec), nat.ptr(F.par), getattr(self.g, "gdiv", 0), F.B,
                     nat.stream_handle(self.device))
            return
        nat.call("mnpbem_bem_solve_batched", F.dtype, n, P, nat.ptr(F.GG[0]), nat.ptr(F.piv[0]),
                 nat.ptr(F.GG[1]), nat.ptr(F.piv[1]), nat.ptr(F.HH[0]), nat.ptr(F.Et), nat.ptr(F.HH[1]),
                 nat.ptr(F.piv[3]), n, nn, n, nat.ptr(F.DG[0]), nat.ptr(F.DG[1]), nat.ptr(F.DS), F.sD,
                 nat.ptr(R), nat.ptr(W), nat.ptr(X), 8 * P, n * 8 * P,
                 nat.ptr(self.g.nvec), nat.ptr(F.par), getattr(self.g, "gdiv", 0), F.B,
                 nat.stream_handle(self.device))

    def solve_batch(self, F, R, fixed_iters=None):
        """Solve A X = R for the batch (R: (B, n, 8P) complex128 on device).
        Returns X (B, n, 8P) complex128.  With ``fixed_iters`` the mixed path
        runs exactly that many direct solves (refinement steps) without
        reading anything back; the last residual norms stay on the device in
        ``self.last_norms`` ((B, 2): ||R - A X||^2, ||R||^2)."""
        B, n = F.B, F.n
        P = R.shape[2] // 8
        dev = self.device
        st = nat.stream_handle(dev)
        if F.dtype == nat.DTYPE_C128:
            X = torch.empty_like(R)
            W = torch.empty_like(R)
            self._solve_direct(F, R, X, W, P)
            self.last_refine_iters = 0
            return X
        # mixed: complex64 direct solve + FP64 refinement
        R32 = torch.empty(R.shape, dtype=torch.complex64, device=dev)
        X32 = torch.empty_like(R32)
        W32 = torch.empty_like(R32)
        X = torch.zeros_like(R)
        Rres = torch.empty_like(R)
        Y = torch.empty((B, n, 16 * P), dtype=torch.complex128, device=dev)
        nrm = torch.empty((B, 2), dtype=torch.float64, device=dev)
        cnt = 2 * R.numel()
        src = R
        if fixed_iters is not None:
            for _ in range(int(fixed_iters)):
                nat.call("mnpbem_convert", nat.DTYPE_C64, cnt, nat.ptr(src), nat.ptr(R32), st)
                self._solve_direct(F, R32, X32, W32, P)
                nat.call("mnpbem_accumulate", nat.DTYPE_C64, cnt, nat.ptr(X32), nat.ptr(X), st)
                self._residual(F, R, X, Rres, Y, nrm, P)
                src = Rres
            self.last_norms = nrm
            return X
        mon = RefinementMonitor(self.refine_tol, self.refine_maxit)
        Xbest = None
        while True:
            nat.call("mnpbem_convert", nat.DTYPE_C64, cnt, nat.ptr(src), nat.ptr(R32), st)
            self._solve_direct(F, R32, X32, W32, P)
            nat.call("mnpbem_accumulate", nat.DTYPE_C64, cnt, nat.ptr(X32), nat.ptr(X), st)
            self._residual(F, R, X, Rres, Y, nrm, P)
            h = nrm.cpu().numpy()
            improved, done = mon.step(np.sqrt(h[:, 0] / np.maximum(h[:, 1], 1e-300)))
            Xbest = _keep_best(Xbest, X, improved)
            if done:
                break
            src = Rres
        self.last_refine_iters = mon.it
        self.last_residual = mon.best
        if F.npflag is not None and np.any(mon.best > mon.tol):
            # Sigma factored without interchanges was not good enough for the
            # refinement: factor the batch again with partial pivoting
            return self.solve_batch(self.factor(F.enei, interchanges=True), R)
        mon.check("BEMRet.solve")
        return Xbest

    def green64(self, F):
        """complex128 transposed G1, H1, G2, H2 of the batch, each (2, B, n, n)
        stacked as (G, H): the copies kept by the mixed path, else assembled
        (the complex128 path factors its copies in place)."""
        if getattr(F, "G64", None) is not None:
            return F.G64, F.H64
        g, n, B = self.g, F.n, F.B
        G = torch.empty((2, B, n, n), dtype=torch.complex128, device=self.device)
        H = torch.empty_like(G)
        for s, hs in ((0, 1.0), (1, -1.0)):
            g.assemble(s, F.kmat, G[s], H[s], hs, nat.DTYPE_C128)
        return G, H

    def solve_adjoint_batch(self, F, C):
        """Solve A^T Y = C for the batch with the factors of ``factor``
        (C: (B, n, 8P) complex128 in the unknowns' layout [sig1 | h1 | sig2 |
        h2]); returns Y (B, n, 8P) in the right-side layout [phi | a | De |
        alpha], so that Y . R = C . X for A X = R.  The mixed path refines in
        FP64 against the complex128 transposed operator like ``solve_batch``."""
        B, n = F.B, F.n
        P = C.shape[2] // 8
        dev = self.device
        st = nat.stream_handle(dev)
        nn = n * n

        def direct(Cin, Yout, Wk):
            if getattr(F, "inverted", False):
                nat.call("mnpbem_bem_solve_adjoint_inv_batched", F.dtype, n, P, nat.ptr(F.GG[0]), nat.ptr(F.GG[1]),
                         nat.ptr(F.HH[0]), nat.ptr(F.Et), nat.ptr(F.HH[1]), n, nn, nat.ptr(Cin), nat.ptr(Wk),
                         nat.ptr(Yout), 8 * P, n * 8 * P, nat.ptr(self.g.nvec), nat.ptr(F.par), B, st)
                return
            nat.call("mnpbem_bem_solve_adjoint_batched", F.dtype, n, P, nat.ptr(F.GG[0]), nat.ptr(F.piv[0]),
                     nat.ptr(F.GG[1]), nat.ptr(F.piv[1]), nat.ptr(F.HH[0]), nat.ptr(F.Et), nat.ptr(F.HH[1]),
                     nat.ptr(F.piv[3]), n, nn, n, nat.ptr(F.DG[0]), nat.ptr(F.DG[1]), nat.ptr(F.DS), F.sD,
                     nat.ptr(Cin), nat.ptr(Wk), nat.ptr(Yout), 8 * P, n * 8 * P, nat.ptr(self.g.nvec),
                     nat.ptr(F.par), B, st)

        if F.dtype == nat.DTYPE_C128:
            Y = torch.empty_like(C)
            direct(C, Y, torch.empty_like(C))
            return Y
        G64, H64 = self.green64(F)
        C32 = torch.empty(C.shape, dtype=torch.complex64, device=dev)
        Y32 = torch.empty_like(C32)
        W32 = torch.empty_like(C32)
        Y = torch.zeros_like(C)
        Cres = torch.empty_like(C)
        Z = torch.empty_like(C)
        T = torch.empty((B, n, 16 * P), dtype=torch.complex128, device=dev)
        nrm = torch.empty((B, 2), dtype=torch.float64, device=dev)
        cnt = 2 * C.numel()
        src = C
        mon = RefinementMonitor(self.refine_tol, self.refine_maxit)
        Ybest = None
        while True:
            nat.call("mnpbem_convert", nat.DTYPE_C64, cnt, nat.ptr(src), nat.ptr(C32), st)
            direct(C32, Y32, W32)
            nat.call("mnpbem_accumulate", nat.DTYPE_C64, cnt, nat.ptr(Y32), nat.ptr(Y), st)
            nat.call("mnpbem_bem_residual_adjoint_batched", n, P, nat.ptr(G64[0]), nat.ptr(H64[0]),
                     nat.ptr(G64[1]), nat.ptr(H64[1]), n, nn, nat.ptr(C), nat.ptr(Y), nat.ptr(Cres), 8 * P,
                     n * 8 * P, nat.ptr(Z), nat.ptr(T), 16 * P, n * 16 * P, nat.ptr(self.g.nvec), nat.ptr(F.par),
                     nat.ptr(nrm), B, st)
            h = nrm.cpu().numpy()
            improved, done = mon.step(np.sqrt(h[:, 0] / np.maximum(h[:, 1], 1e-300)))
            Ybest = _keep_best(Ybest, Y, improved)
            if done:
                break
            src = Cres
        self.last_adjoint_iters = mon.it
        self.last_adjoint_residual = mon.best
        mon.check("BEMRet adjoint solve")
        return Ybest

    def apply_transpose(self, F, Y):
        """A^T Y in complex128 (B, n, 8P), from the complex128 Green matrices."""
        B, n = F.B, F.n
        P = Y.shape[2] // 8
        G64, H64 = self.green64(F)
        out = torch.empty_like(Y)
        Z = torch.empty_like(Y)
        T = torch.empty((B, n, 16 * P), dtype=torch.complex128, device=self.device)
        nat.call("mnpbem_bem_residual_adjoint_batched", n, P, nat.ptr(G64[0]), nat.ptr(H64[0]), nat.ptr(G64[1]),
                 nat.ptr(H64[1]), n, n * n, None, nat.ptr(Y), nat.ptr(out), 8 * P, n * 8 * P, nat.ptr(Z),
                 nat.ptr(T), 16 * P, n * 16 * P, nat.ptr(self.g.nvec), nat.ptr(F.par), None, B,
                 nat.stream_handle(self.device))
        return out

    def apply(self, F, X):
        """A X in complex128 (B, n, 8P), from the complex128 Green matrices."""
        B, n = F.B, F.n
        P = X.shape[2] // 8
        G64, H64 = self.green64(F)
        out = torch.empty_like(X)
        Yw = torch.empty((B, n, 16 * P), dtype=torch.complex128, device=self.device)
        nat.call("mnpbem_bem_residual_batched", n, P, nat.ptr(G64[0]), nat.ptr(H64[0]), nat.ptr(G64[1]),
                 nat.ptr(H64[1]), n, n * n, None, nat.ptr(X), nat.ptr(out), 8 * P, n * 8 * P, nat.ptr(Yw), 16 * P,
                 n * 16 * P, nat.ptr(self.g.nvec), nat.ptr(F.par), None, getattr(self.g, "gdiv", 0), B,
                 nat.stream_handle(self.device))
        return out

    def _residual(self, F, R, X, Rres, Y, nrm, P):
        """Rres = R - A X in complex128 and the squared norms per wavelength;
        Y = [G1 X1 | H1 X1 | G2 X2 | H2 X2] work space."""
        g, n, B = self.g, F.n, F.B
        st = nat.stream_handle(self.device)
        if F.G64 is not None:
            nat.call("mnpbem_bem_residual_batched", n, P, nat.ptr(F.G64[0]), nat.ptr(F.H64[0]),
                     nat.ptr(F.G64[1]), nat.ptr(F.H64[1]), n, n * n, nat.ptr(R), nat.ptr(X), nat.ptr(Rres),
                     8 * P, n * 8 * P, nat.ptr(Y), 16 * P, n * 16 * P, nat.ptr(g.nvec), nat.ptr(F.par),
                     nat.ptr(nrm), getattr(g, "gdiv", 0), B, st)
            return
        if getattr(g, "gdiv", 0):
            raise NotImplementedError("BEMRet: the matrix-free residual needs one geometry per batch")
        P4 = 4 * P
        for s, hs in ((0, 1.0), (1, -1.0)):
            nat.call("mnpbem_green_apply", n, P4, hs, nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(g.area),
                     nat.ptr(g.mat[s]), nat.ptr(F.kmat), g.nmat, g.order, nat.ptr(g._rowptr()), nat.ptr(g.pcol),
                     nat.ptr(g.g), nat.ptr(g.f), nat.ptr(X), 8 * P, n * 8 * P, s * P4, nat.ptr(Y), 16 * P,
                     n * 16 * P, 2 * s * P4, 2 * s * P4 + P4, B, st)
        nat.call("mnpbem_bem_combine_batched", n, P, nat.ptr(Y), 16 * P, n * 16 * P, nat.ptr(R), nat.ptr(Rres),
                 8 * P, n * 8 * P, nat.ptr(g.nvec), nat.ptr(F.par), nat.ptr(nrm), 0, B, st)

    # ------------------------------------------------------------------
    def _state_key(self, enei):
        """What the factors of one wavelength depend on besides the mesh: the
        wavelength and every material's permittivity there (so that a
        material changed in place, e.g. by a gradient step, is refactored)."""
        return (float(enei),) + tuple(complex(e(enei)[0]) for e in self.p.eps)

    def init(self, enei):
        key = self._state_key(enei)
        if self._fac is not None and self._key == key:
            return self
        self._fac = self.factor([enei])
        self.enei = enei
        self._key = key
        return self

    def __call__(self, enei):
        return self.init(enei)

    def solve(self, exc):
        """Surface charges and currents sig1, sig2, h1, h2 (bem_ret.py:1569-1831)
        for an excitation CompStruct: the one PlaneWaveRet.potential returns,
        or any with potentials phi1, phi1p, a1, a1p, phi2, phi2p, a2, a2p
        (missing fields are 0; phi (n,) or (n, npol), a (n, 3) or (n, 3, npol),
        numpy or torch)."""
        enei = exc.enei
        self.init(enei)
        R = excitation_rhs(self.p, self.g, exc)
        X = self.solve_batch(self._fac, R)
        return _unpack(self.p, enei, X[0], R.shape[2] // 8), self

    def __truediv__(self, exc):
        return self.solve(exc)[0]

    def potential(self, sig, inout=2):
        """Potentials at the faces from a solution (CompGreenRet.potential)."""
        return self.g.potential(sig, inout)

    def field(self, sig, inout=2):
        """Fields at the faces from a solution, by the Green-function route
        of CompGreenRet.field (compgreen_ret.py:912-969)."""
        return self.g.field(sig, inout)

    def solve_sweep(self, enei_list, planewave, chunk=16, with_x=False, with_sca=False):
        """Extinction spectrum (n_lambda, npol) of a plane-wave sweep, factored
        and solved ``chunk`` wavelengths per batched launch; with ``with_sca``
        also the scattering spectrum (absorption = ext - sca)."""
        enei_list = np.atleast_1d(np.asarray(enei_list, dtype=float))
        out, scas, xs = [], [], []
        for c0 in range(0, len(enei_list), chunk):
            lam = enei_list[c0:c0 + chunk]
            F = self.factor(lam)
            R = planewave.rhs(self.p, lam, self.g)
            X = self.solve_batch(F, R)
            out.append(planewave.extinction_batch(self.p, lam, X, self.g))
            if with_sca:
                scas.append(planewave.scattering_batch(self.p, lam, X, self.g)[0])
            if with_x:
                xs.append(X)
            del F
        res = [torch.cat(out).cpu().numpy()]
        if with_sca:
            res.append(torch.cat(scas).cpu().numpy())
        if with_x:
            res.append(torch.cat(xs))
        return res[0] if len(res) == 1 else tuple(res)


def needs_interchanges(npflag):
    """Whether a factorisation without row interchanges must be redone with
    them: a multiplier of G1, G2 or Delta above 1 in |re| + |im| (partial
    pivoting would interchange rows), or one of Sigma above the growth guard
    (records of mnpbem_getrf_nopiv_batched, rows G1, G2, Delta, Sigma)."""
    if npflag is None:
        return False
    return int(npflag[:3].max()) > 0 or int(npflag[3].max()) > 1


class SweepGraph:
    """One wavelength-chunk step of the mixed-precision sweep (Green assembly,
    factorisation, right side, ``iters`` refinement steps, extinction)
    captured once as a HIP graph and replayed for any ``B`` wavelengths.

    The step reads its per-wavelength inputs (wavenumbers, permittivities,
    plane-wave parameters) from device buffers that ``run`` fills before each
    replay, and performs no host transfer, allocation outside the graph's
    memory pool or device query while captured; every tensor whose address the
    graph holds is owned by the graph's pool, the solver or this object.  The
    refinement runs a fixed
    number of steps; ``run`` reads the final residual norms after the replay
    and raises if any wavelength misses ``bem.refine_tol`` (more ``iters`` are
    then needed).  The first ``run`` captures; later calls only replay."""

    def __init__(self, bem, planewave, B, iters=3):
        if bem.precision == "double":
            raise ValueError("SweepGraph: the captured step is the mixed-precision path")
        self.bem, self.pw, self.B, self.iters = bem, planewave, int(B), int(iters)
        g, dev = bem.g, bem.device
        self.kmat = torch.empty((self.B, g.nmat, 2), dtype=torch.float64, device=dev)
        self.par = torch.empty((self.B, 5), dtype=torch.float64, device=dev)
        self.pwt = torch.empty((self.B, 7), dtype=torch.float64, device=dev)
        self.store = bem._store_copies(self.B, g.n)
        self.graph = None
        self.captures = 0
        self.fallbacks = 0

    def _inputs(self, lams):
        bem, g = self.bem, self.bem.g
        e_in, e_out = bem._eps_sides(lams)
        self.kmat.copy_(g.wavenumbers(lams))
        self.par.copy_(_par_tensor(e_in, e_out, 2.0 * np.pi / lams, bem.device))
        self.pwt.copy_(self.pw._pw(bem.p, lams, bem.device))

    def _step(self, lams):
        bem, p, g = self.bem, self.bem.p, self.bem.g
        # the graph holds the addresses of the plane wave's side flags and
        # directions; keep them alive (the plane wave's cache may drop them)
        self._keep = self.pw._side_flags(p, g)
        F = bem.factor(lams, kmat=self.kmat, par=self.par, store=self.store)
        R = self.pw.rhs(p, lams, g, pw=self.pwt)
        X = bem.solve_batch(F, R, fixed_iters=self.iters)
        self.ext = self.pw.extinction_batch(p, lams, X, g, pw=self.pwt)
        self.norms = bem.last_norms
        self.npflag = F.npflag

    def _eager(self, lams):
        """The chunk by the eager path with partial pivoting throughout."""
        bem = self.bem
        F = bem.factor(lams, interchanges=True)
        X = bem.solve_batch(F, self.pw.rhs(bem.p, lams, bem.g))
        self.last_residual = bem.last_residual
        self.fallbacks += 1
        return self.pw.extinction_batch(bem.p, lams, X, bem.g).cpu().numpy()

    def run(self, enei_list):
        lams = np.atleast_1d(np.asarray(enei_list, dtype=float))
        if len(lams) != self.B:
            raise ValueError("SweepGraph: captured for %d wavelengths, got %d" % (self.B, len(lams)))
        self._inputs(lams)
        if self.graph is None:
            dev = self.bem.device
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                self._step(lams)   # eager warm-up on a side stream (library, caches)
            torch.cuda.current_stream(dev).wait_stream(side)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._step(lams)
            self.captures += 1
        self.graph.replay()
        if needs_interchanges(self.npflag):
            # the captured factorisation ran without the row interchanges
            # that these wavelengths need: solve them eagerly with them
            return self._eager(lams)
        h = self.norms.cpu().numpy()
        rel = np.sqrt(h[:, 0] / np.maximum(h[:, 1], 1e-300))
        self.last_residual = rel
        if np.any(rel > self.bem.refine_tol) and self.npflag is not None:
            # Sigma without interchanges needs more steps than captured
            return self._eager(lams)
        if np.any(rel > self.bem.refine_tol):
            raise RuntimeError("SweepGraph: residual %.2e above refine_tol after %d steps; use more iters"
                               % (rel.max(), self.iters))
        return self.ext.cpu().numpy()


_POTENTIALS = ("phi1", "phi1p", "a1", "a1p", "phi2", "phi2p", "a2", "a2p")


def excitation_rhs(p, g, exc):
    """Right side R (1, n, 8P) = [phi | a | De | alpha] of the BEM equations
    for an excitation CompStruct (BEMRet._excitation, bem_ret.py:1360-1429):
    the plane wave's precomputed right side, or the one the HIP kernel forms
    from the potentials phi1, phi1p, a1, a1p, phi2, phi2p, a2, a2p."""
    R = exc.get("_R")
    if R is not None:
        return R
    n, dev = p.n, g.device
    vals, P = {}, None
    for name in _POTENTIALS:
        v = exc.get(name)
        if v is None or (np.isscalar(v) and v == 0):
            vals[name] = None
            continue
        t = (v if torch.is_tensor(v) else torch.as_tensor(np.asarray(v))).to(device=dev, dtype=torch.complex128)
        if name.startswith("phi"):
            if t.ndim not in (1, 2) or t.shape[0] != n:
                raise ValueError("BEMRet.solve: %s must have shape (n,) or (n, npol) with n = %d" % (name, n))
            t = t.reshape(n, -1)
            q = t.shape[1]
        else:
            if t.ndim not in (2, 3) or tuple(t.shape[:2]) != (n, 3):
                raise ValueError("BEMRet.solve: %s must have shape (n, 3) or (n, 3, npol) with n = %d" % (name, n))
            t = t.reshape(n, 3, -1)
            q = t.shape[2]
        if P is not None and q != P:
            raise ValueError("BEMRet.solve: the potentials have different numbers of polarisations")
        P = q
        vals[name] = t.contiguous()
    if P is None:
        raise ValueError("BEMRet.solve: the excitation has no potentials (phi1, a1, ..., a2p)")
    if P > 4:
        raise ValueError("BEMRet.solve: at most 4 polarisations per solve")
    enei = exc.enei
    eps = lambda side: torch.as_tensor(np.array([complex(p.eps[m - 1](enei)[0]) for m in p.inout_faces[:, side]]),
                                       dtype=torch.complex128).to(dev)
    e1, e2 = eps(0), eps(1)
    R = torch.empty((1, n, 8 * P), dtype=torch.complex128, device=dev)
    nat.call("mnpbem_excitation", n, P, *[nat.ptr(vals[k]) for k in _POTENTIALS], nat.ptr(g.nvec), nat.ptr(e1),
             nat.ptr(e2), 2.0 * np.pi / enei, nat.ptr(R), 8 * P, nat.stream_handle(dev))
    return R


class RefinementMonitor:
    """Stopping rule of the FP64 iterative refinement of a batch of systems.

    ``step(rel)`` records the relative residuals ||R - A X|| / ||R|| after one
    more correction and returns (improved, done): which systems reached their
    smallest residual so far (their iterate is the one to keep) and whether to
    stop.  Refinement stops when every system is at or below ``tol``, after
    ``maxit`` + 1 solves, or when no system still above ``tol`` gained at
    least a factor ``stall`` in the last step (it has stagnated or diverges,
    as it can for nearly singular systems).  ``check`` raises when a system
    missed ``tol``."""

    def __init__(self, tol, maxit, stall=0.5):
        self.tol, self.maxit, self.stall = float(tol), int(maxit), float(stall)
        self.best = None
        self.it = 0

    def step(self, rel):
        rel = np.asarray(rel, dtype=float)
        self.it += 1
        if self.best is None:
            self.best = rel.copy()
            return np.ones(rel.shape, dtype=bool), bool(np.all(rel <= self.tol)) or self.it > self.maxit
        improved = rel < self.best
        still_open = self.best > self.tol
        gaining = np.any(rel[still_open] < self.stall * self.best[still_open])
        self.best = np.where(improved, rel, self.best)
        done = bool(np.all(self.best <= self.tol)) or self.it > self.maxit or not gaining
        return improved, done

    def check(self, what):
        if self.best is not None and np.any(self.best > self.tol):
            raise RuntimeError("%s: mixed-precision refinement stopped at relative residual %.2e > refine_tol "
                               "%.1e after %d solves (the system is too ill-conditioned for the complex64 "
                               "factors; use precision='double')" % (what, float(np.max(self.best)), self.tol,
                                                                     self.it))


def _keep_best(best, X, improved):
    """The iterate with the smallest residual per system (batch index 0)."""
    if best is None or np.all(improved):
        return X.clone() if best is None else best.copy_(X)
    idx = np.nonzero(improved)[0]
    if len(idx):
        t = torch.as_tensor(idx, device=X.device)
        best[t] = X[t]
    return best


def _unpack(p, enei, X, P):
    n = X.shape[0]
    sig1 = X[:, 0:P]
    h1 = X[:, P:4 * P].reshape(n, 3, P)
    sig2 = X[:, 4 * P:5 * P]
    h2 = X[:, 5 * P:8 * P].reshape(n, 3, P)
    if P == 1:
        sig1, sig2, h1, h2 = sig1[:, 0], sig2[:, 0], h1[:, :, 0], h2[:, :, 0]
    return CompStruct(p, enei, sig1=sig1, sig2=sig2, h1=h1, h2=h2, _X=X[None])
