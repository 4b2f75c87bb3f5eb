"""Iterative retarded BEM solver: GMRES on the GPU.

``BEMRetIter`` keeps the interface of mnpbem/bem/bem_ret_iter.py and
mnpbem/bem/bem_iter.py (solver, tol, maxit, restart, precond, hmatrix, htol,
kmax, cleaf, fadmiss, eta, preconditioner; ``solve(exc) -> (sig, self)``,
``info()``).  The unknowns are the packed 8N vector [sig1 | h1 | sig2 | h2]
and the right side [phi | a | De | alpha] of GdA&H Eqs. (10, 11, 14, 17)
(bem_ret_iter.py:1196-1232, 1362-1501).

Per batch of wavelengths the operator is applied in FP64 from complex128
Green matrices (dense) or from their H-matrices (``hmatrix=True``: ACA on the
GPU, blocks stored in ``hmat_dtype``), and the preconditioner is the direct
BEMRet solve (bem_ret_iter.py:1503-1634 applies the same Sigma / Delta
factorisation), in complex128 (``precision='double'``) or complex64 on
matrix cores (``'mixed'``, ``'mixed-fp32'``).  GMRES is flexible and right
preconditioned, restarted every ``restart`` iterations, with classical
Gram-Schmidt applied twice; each (wavelength, polarisation) is its own
system, all advancing in lockstep, and a system has converged when its true
residual satisfies ||b - A x|| <= tol ||b|| (the reference stacks the
polarisations into one vector and applies the same test to the stack).
"""

import numpy as np
import torch

from . import _native as nat
from .bem import BEMRet, _unpack, excitation_rhs
from .greenfun import CompGreenRet


class BEMIter:
    """Options and statistics of the iterative solvers (bem_iter.py:11-244)."""

    SOLVER_MAP = {"gmres": "gmres", "cgs": "cgs", "bicgstab": "bicgstab"}

    def __init__(self, solver="gmres", tol=1e-4, maxit=200, restart=None, precond="hmat", output=0, **kwargs):
        self.solver = solver
        self.tol = tol
        self.maxit = maxit
        self.restart = restart
        self.precond = precond
        self.output = output
        self._flag = None
        self._relres = None
        self._iter = None
        self._stat = None

    def _set_iter(self, flag, relres, iter_info):
        if self._flag is None:
            self._flag, self._relres, self._iter = [], [], []
        self._flag.append(flag)
        self._relres.append(relres)
        self._iter.append(np.array(iter_info).copy())

    def info(self):
        """(flags, relative residuals, [restart cycles, iterations]) of every solve."""
        return self._flag, self._relres, self._iter

    def _print_stat(self, flag, relres, iter_info):
        print("gmres(%s), it=%3d(%d), res=%10.4g, flag=%d" % (self.maxit, int(iter_info[1]), int(iter_info[0]),
                                                              relres, flag))

    @staticmethod
    def options(**kwargs):
        op = {"solver": "gmres", "tol": 1e-6, "maxit": 100, "restart": None, "precond": "hmat", "output": 0,
              "cleaf": 200, "htol": 1e-6, "kmax": [4, 100],
              "fadmiss": lambda rad1, rad2, dist: 2.5 * min(rad1, rad2) < dist}
        for k, v in kwargs.items():
            if k in op:
                op[k] = v
        return op

    def __repr__(self):
        return "BEMIter(solver={}, tol={}, maxit={}, precond={})".format(self.solver, self.tol, self.maxit,
                                                                        self.precond)


class _Batch:
    """Operator and preconditioner state of one batch of wavelengths."""


class BEMRetIter(BEMIter):
    """Retarded BEM solver with GMRES on the GPU."""

    name = "bemsolver"
    needs = {"sim": "ret"}

    def __init__(self, p, enei=None, precision="double", green=None, hmat_dtype="complex128", **options):
        self._hmatrix = bool(options.pop("hmatrix", False))
        self._htol = options.pop("htol", 1e-6)
        self._kmax = options.pop("kmax", [4, 100])
        self._cleaf = options.pop("cleaf", 200)
        self._fadmiss = options.pop("fadmiss", None)
        self._eta = options.pop("eta", 2.5)
        self._hlu_mode = options.pop("preconditioner", "auto")
        self._htol_precond = options.pop("htol_precond", 1e-4)
        # block preconditioner: apply explicit inverses of G1, G2 and Sigma
        # (matrix-vector products) instead of their triangular factors
        self._precond_inverse = bool(options.pop("precond_inverse", True))
        for unsupported in ("schur", "refun"):
            if options.pop(unsupported, None):
                raise NotImplementedError("BEMRetIter: option '%s' is not provided" % unsupported)
        if self._hmatrix and "precond" not in options:
            options["precond"] = None
        if self._hlu_mode not in ("auto", "none", "hlu_dense", "hlu_tree", "block"):
            raise ValueError("BEMRetIter: unknown preconditioner '%s'" % self._hlu_mode)
        if self._hlu_mode == "auto" and self._hmatrix and len(getattr(p, "p", [])) > 1:
            # H-matrix operator over several particles: the per-particle
            # direct solves, so that no matrix spans all faces
            self._hlu_mode = "block"
        super().__init__(**options)
        if self.solver not in self.SOLVER_MAP:
            raise ValueError("[error] iterative solver not known: <{}>".format(self.solver))
        if self.solver != "gmres":
            raise NotImplementedError("BEMRetIter: only solver='gmres' runs on the GPU")
        self.p = p
        self.precision = precision
        self.hmat_dtype = hmat_dtype
        self.g = green if green is not None else CompGreenRet(p, p)
        self.device = self.g.device
        self._direct = BEMRet(p, precision=precision, green=self.g)
        # preconditioner='block': the direct solve of each particle on its own
        # faces (the diagonal blocks of the 8N operator: a particle's Green
        # functions, refinement and closed-surface correction do not involve
        # the other particles); GMRES carries the coupling between particles
        # Particles that are translated copies of an earlier one (same faces
        # and materials, nodes equal up to a shift) reuse its solver: their
        # diagonal blocks agree to rounding, which GMRES absorbs.
        self._blocks = None
        self._block_bems = None
        if self._hlu_mode == "block":
            from .geometry import ComParticle
            self._blocks, self._block_bems, firsts = [], [], []
            for k, part in enumerate(p.p):
                u = next((q for q, (kk, _) in enumerate(firsts)
                          if _translated_copy(p.p[kk], p.inout[kk], part, p.inout[k])), None)
                if u is None:
                    u = len(self._block_bems)
                    firsts.append((k, part))
                    sub = ComParticle(p.eps, [part], [list(p.inout[k])], 1)
                    self._block_bems.append(BEMRet(sub, precision=precision, residual="matrix_free"))
                self._blocks.append((int(p.offsets[k]), u))
        self._aca = None
        if self._hmatrix:
            from .hmatrix import ACACompGreenRet
            kmax = max(self._kmax) if hasattr(self._kmax, "__iter__") else self._kmax
            htol = max(self._htol) if hasattr(self._htol, "__iter__") else self._htol
            self._aca = ACACompGreenRet(p, htol=htol, kmax=kmax, cleaf=self._cleaf, fadmiss=self._fadmiss,
                                        eta=self._eta, dtype=hmat_dtype, green=self.g)
        self.enei = None
        self._state = None
        self.last_hmats = None
        if enei is not None:
            self.init(enei)

    # ------------------------------------------------------------------
    @property
    def _use_precond(self):
        if self._hmatrix:
            return self._hlu_mode != "none" or self.precond is not None
        return self.precond is not None

    def _prepare(self, enei_list):
        """Operator and preconditioner for a batch of wavelengths."""
        enei_list = np.atleast_1d(np.asarray(enei_list, dtype=float))
        g, dev = self.g, self.device
        B, n = len(enei_list), g.n
        S = _Batch()
        S.enei, S.B, S.n = enei_list, B, n
        S.F = None
        S.Fk = None
        if self._blocks is not None:
            S.Fk = [bem.factor(enei_list) for bem in self._block_bems]
            if self._precond_inverse:
                # GMRES applies the block factorisations once per iteration to
                # a few right-hand sides: explicit inverses turn the
                # latency-bound triangular steps into matrix-vector products
                for F, bem in zip(S.Fk, self._block_bems):
                    bem.invert_factors(F)
        if self._use_precond and self._blocks is None:
            S.F = self._direct.factor(enei_list)
            par = S.F.par
        else:
            e_in, e_out = self._direct._eps_sides(enei_list)
            from .bem import _par_tensor
            par = _par_tensor(e_in, e_out, 2.0 * np.pi / enei_list, dev)
        S.par = par
        if self._hmatrix:
            # H-matrices of G1, H1, G2, H2 per wavelength, entries generated
            # in place by the ACA kernels (no n x n matrix); GMRES only
            # multiplies with them, so each keeps its packed product layouts
            # and frees its block storage before the next is generated
            kmat = g.wavenumbers(enei_list)
            S.hm = [[self._aca.generate(s, key, kmat[b], lam).release_blocks()
                     for s, key in ((0, "G"), (0, "H1"), (1, "G"), (1, "H2"))]
                    for b, lam in enumerate(enei_list)]
            S.ops = None
            if S.F is not None:
                S.F.G64 = S.F.H64 = None
            self.last_hmats = S.hm
            return S
        if S.F is not None and getattr(S.F, "G64", None) is not None:
            ops = (S.F.G64[0], S.F.H64[0], S.F.G64[1], S.F.H64[1])
        else:
            kmat = g.wavenumbers(enei_list)
            G = torch.empty((2, B, n, n), dtype=torch.complex128, device=dev)
            H = torch.empty((2, B, n, n), dtype=torch.complex128, device=dev)
            for s, hs in ((0, 1.0), (1, -1.0)):
                g.assemble(s, kmat, G[s], H[s], hs, nat.DTYPE_C128)
            ops = (G[0], H[0], G[1], H[1])
        S.ops = ops
        S.hm = None
        return S

    def _apply(self, S, X, Out, Y, P):
        """Out = A X (8N operator, complex128) for the batch."""
        n, B = S.n, S.B
        st = nat.stream_handle(self.device)
        if S.hm is None:
            G1, H1, G2, H2 = S.ops
            nat.call("mnpbem_bem_residual_batched", n, P, nat.ptr(G1), nat.ptr(H1), nat.ptr(G2), nat.ptr(H2), n,
                     n * n, None, nat.ptr(X), nat.ptr(Out), 8 * P, n * 8 * P, nat.ptr(Y), 16 * P, n * 16 * P,
                     nat.ptr(self.g.nvec), nat.ptr(S.par), None, 0, B, st)
            return
        P4 = 4 * P
        for b in range(B):
            xo = b * n * 8 * P
            yo = b * n * 16 * P
            for q, hm in enumerate(S.hm[b]):
                hm.matvec_device(nat.ptr_off(X, xo + (0 if q < 2 else P4)), 8 * P, nat.ptr_off(Y, yo + q * P4),
                                 16 * P, P4, stream=st)
        nat.call("mnpbem_bem_combine_batched", n, P, nat.ptr(Y), 16 * P, n * 16 * P, None, nat.ptr(Out), 8 * P,
                 n * 8 * P, nat.ptr(self.g.nvec), nat.ptr(S.par), None, 0, B, st)

    def _precond_blocks(self, S, V, Z, ws, P):
        """Z = blockdiag(A_kk)^-1 V: each particle's direct solve on its rows
        of the packed vectors (row stride 8P, wavelength stride n 8P).  The
        particles sharing one factorisation at equally spaced offsets (the
        translated copies) are solved as one batch whose systems read the
        same factors (factor strides 0) and whose vectors are the copies'
        row ranges."""
        n, B = S.n, S.B
        st = nat.stream_handle(self.device)
        c128 = S.Fk[0].dtype == nat.DTYPE_C128
        if c128:
            R, X, W = V, Z, ws["W"]
        else:
            R, X, W = ws["R32"], ws["X32"], ws["W32"]
            nat.call("mnpbem_convert", nat.DTYPE_C64, 2 * V.numel(), nat.ptr(V), nat.ptr(R), st)
        for u, offs, step, par in self._block_groups(S):
            F, bem = S.Fk[u], self._block_bems[u]
            nn = F.n * F.n
            cnt = len(offs)
            for b in range(B):
                o = (b * n + offs[0]) * 8 * P
                fo = lambda t, stride: nat.ptr_off(t, b * stride)
                if getattr(F, "inverted", False):
                    nat.call("mnpbem_bem_solve_inv_batched", F.dtype, F.n, P, fo(F.GG[0], nn), fo(F.GG[1], nn),
                             fo(F.HH[0], nn), fo(F.Et, nn), fo(F.HH[1], nn), F.n, 0, nat.ptr_off(R, o),
                             nat.ptr_off(W, o), nat.ptr_off(X, o), 8 * P, step * 8 * P, nat.ptr(bem.g.nvec),
                             nat.ptr_off(par, 5 * cnt * b), 0, cnt, st)
                    continue
                nat.call("mnpbem_bem_solve_batched", F.dtype, F.n, P, fo(F.GG[0], nn), fo(F.piv[0], F.n),
                         fo(F.GG[1], nn), fo(F.piv[1], F.n), fo(F.HH[0], nn), fo(F.Et, nn), fo(F.HH[1], nn),
                         fo(F.piv[3], F.n), F.n, 0, 0, fo(F.DG[0], F.sD), fo(F.DG[1], F.sD), fo(F.DS, F.sD), 0,
                         nat.ptr_off(R, o), nat.ptr_off(W, o), nat.ptr_off(X, o), 8 * P, step * 8 * P,
                         nat.ptr(bem.g.nvec), nat.ptr_off(par, 5 * cnt * b), 0, cnt, st)
        if not c128:
            nat.call("mnpbem_widen", nat.DTYPE_C64, 2 * V.numel(), nat.ptr(X), nat.ptr(Z), st)

    def _block_groups(self, S):
        """(factor index, row offsets, offset step, per-system parameters) of
        each batch of block solves; cached on the wavelength batch."""
        groups = getattr(S, "block_groups", None)
        if groups is not None:
            return groups
        groups = []
        for u in range(len(self._block_bems)):
            offs = [off for off, uu in self._blocks if uu == u]
            steps = np.diff(offs)
            if len(offs) > 1 and np.all(steps == steps[0]) and steps[0] >= self._block_bems[u].p.n:
                runs = [offs]
            else:
                runs = [[off] for off in offs]
            for r in runs:
                step = int(r[1] - r[0]) if len(r) > 1 else 0
                par = S.Fk[u].par.reshape(S.B, 1, 5).expand(S.B, len(r), 5).contiguous()
                groups.append((u, r, step, par))
        S.block_groups = groups
        return groups

    def _precond(self, S, V, Z, ws, P):
        """Z = M^-1 V with the direct BEMRet solve (or Z = V without one)."""
        if S.Fk is not None:
            self._precond_blocks(S, V, Z, ws, P)
            return
        if S.F is None:
            Z.copy_(V)
            return
        F = S.F
        if F.dtype == nat.DTYPE_C128:
            self._direct._solve_direct(F, V, Z, ws["W"], P)
            return
        st = nat.stream_handle(self.device)
        cnt = 2 * V.numel()
        nat.call("mnpbem_convert", nat.DTYPE_C64, cnt, nat.ptr(V), nat.ptr(ws["R32"]), st)
        self._direct._solve_direct(F, ws["R32"], ws["X32"], ws["W32"], P)
        nat.call("mnpbem_widen", nat.DTYPE_C64, cnt, nat.ptr(ws["X32"]), nat.ptr(Z), st)

    # ------------------------------------------------------------------
    def _dots(self, k, V, sVk, W, out, work, L, P, B):
        nat.call("mnpbem_mv_dot_batched", L, P, k, nat.ptr(V), sVk, L, nat.ptr(W), L, nat.ptr(out), nat.ptr(work),
                 B, nat.stream_handle(self.device))

    def _gmres(self, S, b):
        """Flexible right-preconditioned GMRES(restart) on all B x P systems."""
        dev = self.device
        st = nat.stream_handle(dev)
        B, n, C = b.shape
        P = C // 8
        L = n * C
        NS = B * P
        m = self.restart if self.restart is not None else min(8 * n * P, self.maxit)
        m = max(1, int(m))
        tol = float(self.tol)
        ws = {"W": torch.empty_like(b)}
        F0 = S.F if S.F is not None else (S.Fk[0] if S.Fk else None)
        if F0 is not None and F0.dtype != nat.DTYPE_C128:
            ws["R32"] = torch.empty(b.shape, dtype=torch.complex64, device=dev)
            ws["X32"] = torch.empty_like(ws["R32"])
            ws["W32"] = torch.empty_like(ws["R32"])
        Y = torch.empty((B, n, 16 * P), dtype=torch.complex128, device=dev)
        V = torch.empty((m + 1, B, n, C), dtype=torch.complex128, device=dev)
        Z = torch.empty((m, B, n, C), dtype=torch.complex128, device=dev)
        work = torch.empty((B * (m + 1) * 64 * P,), dtype=torch.complex128, device=dev)
        hd = torch.empty((B, m + 1, P), dtype=torch.complex128, device=dev)
        hd2 = torch.empty_like(hd)
        nrm = torch.empty((B, 1, P), dtype=torch.complex128, device=dev)
        sVk = B * L

        def norms(vec):
            self._dots(1, vec, 0, vec, nrm, work, L, P, B)
            return np.sqrt(np.maximum(nrm.cpu().numpy().real.reshape(NS), 0.0)), nrm

        bn, _ = norms(b)
        x = torch.zeros_like(b)
        r = torch.empty_like(b)
        cycles = iters = 0
        converged = bn == 0.0
        rel = np.zeros(NS)
        for cycle in range(max(1, int(self.maxit))):
            if cycle == 0:
                r.copy_(b)
            else:
                self._apply(S, x, r, Y, P)
                r.mul_(-1.0).add_(b)
            beta, nr = norms(r)
            rel = np.where(bn > 0, beta / np.where(bn > 0, bn, 1.0), 0.0)
            converged = rel <= tol
            if np.all(converged):
                break
            cycles += 1
            nat.call("mnpbem_mv_scale_batched", L, P, 1, nat.ptr(nr), nat.ptr(r), L, nat.ptr(V[0]), L, B, st)
            H = np.zeros((NS, m + 1, m), dtype=np.complex128)
            cs = np.zeros((NS, m))
            sn = np.zeros((NS, m), dtype=np.complex128)
            gv = np.zeros((NS, m + 1), dtype=np.complex128)
            gv[:, 0] = beta
            kdone = np.zeros(NS, dtype=np.int64)
            live = ~converged
            for j in range(m):
                self._precond(S, V[j], Z[j], ws, P)
                w = V[j + 1]
                self._apply(S, Z[j], w, Y, P)
                self._dots(j + 1, V, sVk, w, hd, work, L, P, B)
                nat.call("mnpbem_mv_update_batched", L, P, j + 1, nat.ptr(V), sVk, L, nat.ptr(hd), -1.0, nat.ptr(w),
                         L, B, st)
                self._dots(j + 1, V, sVk, w, hd2, work, L, P, B)
                nat.call("mnpbem_mv_update_batched", L, P, j + 1, nat.ptr(V), sVk, L, nat.ptr(hd2), -1.0,
                         nat.ptr(w), L, B, st)
                self._dots(1, w, 0, w, nrm, work, L, P, B)
                nat.call("mnpbem_mv_scale_batched", L, P, 1, nat.ptr(nrm), nat.ptr(w), L, nat.ptr(w), L, B, st)
                kk = B * (j + 1) * P  # the dot kernels wrote (B, j+1, P) contiguously
                h = (hd.view(-1)[:kk] + hd2.view(-1)[:kk]).view(B, j + 1, P).cpu().numpy()
                hn = np.sqrt(np.maximum(nrm.cpu().numpy().real, 0.0)).reshape(NS)
                iters += 1
                col = np.transpose(h, (0, 2, 1)).reshape(NS, j + 1)
                H[:, :j + 1, j] = col
                H[:, j + 1, j] = hn
                # apply the earlier rotations, then a new one zeroing H[j+1, j]
                hc = np.concatenate([col, hn[:, None].astype(np.complex128)], axis=1)
                for i in range(j):
                    a, bb = hc[:, i].copy(), hc[:, i + 1].copy()
                    hc[:, i] = cs[:, i] * a + sn[:, i] * bb
                    hc[:, i + 1] = -np.conj(sn[:, i]) * a + cs[:, i] * bb
                a, bb = hc[:, j], hc[:, j + 1]
                ra = np.abs(a)
                rr = np.sqrt(ra ** 2 + np.abs(bb) ** 2)
                safe = rr > 0
                c_new = np.where(safe & (ra > 0), ra / np.where(safe, rr, 1.0), 0.0)
                s_new = np.where(ra > 0, (a / np.where(ra > 0, ra, 1.0)) * np.conj(bb) / np.where(safe, rr, 1.0),
                                 1.0 + 0j)
                cs[:, j], sn[:, j] = c_new, s_new
                hc[:, j] = c_new * a + s_new * bb
                hc[:, j + 1] = 0.0
                H[:, :j + 2, j] = hc
                g0 = gv[:, j].copy()
                gv[:, j] = c_new * g0
                gv[:, j + 1] = -np.conj(s_new) * g0
                est = np.abs(gv[:, j + 1])
                kdone = np.where(live, j + 1, kdone)
                done_now = live & ((est <= tol * bn) | (hn <= 1e-300 * np.maximum(bn, 1e-300)))
                live = live & ~done_now
                if not np.any(live):
                    break
            # least-squares solution per system and x += Z y
            kmax = int(kdone.max())
            yv = np.zeros((NS, kmax), dtype=np.complex128)
            for s in range(NS):
                k = int(kdone[s])
                if k == 0:
                    continue
                R = np.triu(H[s, :k, :k])
                yv[s, :k] = _back_substitute(R, gv[s, :k])
            coef = torch.as_tensor(np.ascontiguousarray(yv.reshape(B, P, kmax).transpose(0, 2, 1))).to(dev)
            nat.call("mnpbem_mv_update_batched", L, P, kmax, nat.ptr(Z), sVk, L, nat.ptr(coef), 1.0, nat.ptr(x), L,
                     B, st)
        else:
            self._apply(S, x, r, Y, P)
            r.mul_(-1.0).add_(b)
            beta, _ = norms(r)
            rel = np.where(bn > 0, beta / np.where(bn > 0, bn, 1.0), 0.0)
            converged = rel <= tol
        flag = 0 if np.all(converged) else int(iters)
        return x, flag, rel, (cycles, iters)

    # ------------------------------------------------------------------
    def init(self, enei):
        # keyed on the wavelength and every material's permittivity there, so
        # that a material changed in place is prepared again
        key = (float(enei),) + tuple(complex(e(enei)[0]) for e in self.p.eps)
        if getattr(self, "_state", None) is not None and getattr(self, "_key", None) == key:
            return self
        self._state = self._prepare([enei])
        self.enei = enei
        self._key = key
        return self

    def __call__(self, enei):
        return self.init(enei)

    def solve_batch(self, S, R):
        """X solving A X = R for the batch (R (B, n, 8P) complex128)."""
        P = R.shape[2] // 8
        if self.maxit == 0 or self.solver is None:
            if S.F is None and S.Fk is None:
                raise ValueError("[error] preconditioner must be hmat when maxit=0")
            X = torch.empty_like(R)
            self._precond(S, R, X, {"W": torch.empty_like(R), "R32": torch.empty(R.shape, dtype=torch.complex64,
                                                                                  device=R.device),
                                    "X32": torch.empty(R.shape, dtype=torch.complex64, device=R.device),
                                    "W32": torch.empty(R.shape, dtype=torch.complex64, device=R.device)}, P)
            return X
        X, flag, rel, it = self._gmres(S, R)
        self._set_iter(flag, float(np.max(rel)) if len(rel) else 0.0, it)
        if self.output:
            self._print_stat(flag, float(np.max(rel)), it)
        return X

    def solve(self, exc):
        """Surface charges and currents (bem_ret_iter.py:1636-1700) for an
        excitation CompStruct: a plane wave's, or any with potentials phi1,
        phi1p, a1, a1p, phi2, phi2p, a2, a2p (as BEMRet.solve)."""
        enei = exc.enei
        self.init(enei)
        R = excitation_rhs(self.p, self.g, exc)
        X = self.solve_batch(self._state, R)
        return _unpack(self.p, enei, X[0], R.shape[2] // 8), self

    def __truediv__(self, exc):
        return self.solve(exc)[0]

    def potential(self, sig, inout=2):
        """Potentials at the faces from a solution (bem_ret_iter.py potential)."""
        return self.g.potential(sig, inout)

    def field(self, sig, inout=2):
        """Fields at the faces by the Green-function route of
        CompGreenRet.field (compgreen_ret.py:912-969)."""
        return self.g.field(sig, inout)

    def solve_sweep(self, enei_list, planewave, chunk=8):
        """Extinction (n_lambda, npol) of a plane-wave sweep."""
        enei_list = np.atleast_1d(np.asarray(enei_list, dtype=float))
        out = []
        for c0 in range(0, len(enei_list), chunk):
            lam = enei_list[c0:c0 + chunk]
            S = self._prepare(lam)
            R = planewave.rhs(self.p, lam, self.g)
            X = self.solve_batch(S, R)
            out.append(planewave.extinction_batch(self.p, lam, X, self.g))
            del S
        return torch.cat(out).cpu().numpy()

    def clear(self):
        self._state = None
        self.enei = None
        self._key = None
        return self

    def hinfo(self):
        """Compression of the H-matrices of the last batch."""
        if not self.last_hmats:
            return None
        c = [h.compression() for row in self.last_hmats for h in row]
        print("\nCompression Green functions        :  {:8.6f}".format(float(np.mean(c))))
        return float(np.mean(c))

    def __repr__(self):
        status = "enei={:.1f}nm".format(self.enei) if self.enei is not None else "not initialized"
        return "BEMRetIter(p: {} faces, solver={}, {})".format(self.p.n, self.solver, status)


def _translated_copy(a, io_a, b, io_b, rel=1e-12):
    """True when particle b is particle a shifted: the same element layout,
    materials and interpolation, and nodes equal up to one translation (to
    rel of a's extent)."""
    if type(a) is not type(b) or a.interp != b.interp or a.n != b.n or list(io_a) != list(io_b):
        return False
    if not np.array_equal(a.nvert, b.nvert):
        return False
    used = np.ones(a.nodes9.shape[:2], dtype=bool)   # a triangle leaves slots 3, 7, 8 empty
    used[np.ix_(a.nvert == 3, [3, 7, 8])] = False
    x, y = a.nodes9[used], b.nodes9[used]
    d = y - x
    ext = max(float(np.ptp(x, axis=0).max()), 1e-300)
    return bool(np.abs(d - d[0]).max() <= rel * ext)


def _back_substitute(R, g):
    k = R.shape[0]
    y = np.zeros(k, dtype=np.complex128)
    for i in range(k - 1, -1, -1):
        y[i] = (g[i] - R[i, i + 1:] @ y[i + 1:]) / R[i, i] if R[i, i] != 0 else 0.0
    return y
