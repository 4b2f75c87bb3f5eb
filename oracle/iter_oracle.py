"""FP64 CPU oracle of the iterative retarded BEM solve (BEMRetIter).

TEST INFRASTRUCTURE ONLY (imported by ``tests/`` as the checker).

Restates mnpbem/bem/bem_ret_iter.py: the column-major packing of the 8N
vector (``_pack`` / ``_unpack`` :1196-1232), the operator ``_afun``
(:1362-1501, scalar-eps branch), the preconditioner ``_mfun`` (:1503-1634,
the BEMRet direct solve from the same Sigma / Delta factors) and the call of
scipy.sparse.linalg.gmres in mnpbem/bem/bem_iter.py:79-89 (restart =
min(len(b), maxit) when none is given).

Pinning: the direct solve it preconditions with is pinned by the oracle's
operator identity; tests check that GMRES reaches ||b - A x|| <= tol ||b||
and that the iterative and direct solutions agree, as the reference's own
tests/test_bem_iter.py and tests/test_iter_multipol.py do.
"""

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres

from .bem_oracle import BEMRetOracle, bem_operator_apply, green_matrices


def pack(phi, a, phip, ap):
    """bem_ret_iter.py:1196-1211: column-major flatten of the four parts."""
    return np.concatenate([np.asarray(t).ravel(order="F") for t in (phi, a, phip, ap)])


def unpack(vec, n):
    """bem_ret_iter.py:1213-1232."""
    siz = vec.size // (8 * n)
    v2 = vec.reshape(-1, 8, order="F")
    phi = v2[:, 0].reshape(n, siz, order="F")
    a = v2[:, 1:4].reshape(n, 3, siz, order="F")
    phip = v2[:, 4].reshape(n, siz, order="F")
    ap = v2[:, 5:8].reshape(n, 3, siz, order="F")
    return phi, a, phip, ap


class BEMRetIterOracle:
    """scipy GMRES on the 8N system with the direct solve as preconditioner."""

    def __init__(self, setup, tol=1e-6, maxit=200, restart=None, precond=True):
        self.setup = setup
        self.comp = setup.comp
        self.tol, self.maxit, self.restart, self.precond = tol, maxit, restart, precond
        self.direct = BEMRetOracle(setup)

    def solve(self, exc):
        enei = exc["enei"]
        d = self.direct.init(enei)
        comp = self.comp
        n = comp.n
        k = 2.0 * np.pi / enei
        mats = d.mats
        npol = exc["a2"].shape[2]
        # right side: the same (phi, a, De, alpha) the direct solve builds
        z1 = np.zeros((n, npol), complex)
        e1, e2 = d.eps1, d.eps2
        nv = comp.nvec
        phi = exc.get("phi2", z1) - exc.get("phi1", z1)
        a = exc["a2"] - exc["a1"]
        alpha = (exc["a2p"] - exc["a1p"]) - 1j * k * nv[:, :, None] * (
            exc.get("phi2", z1) * e2 - exc.get("phi1", z1) * e1)[:, None, :]
        De = (e2 * exc.get("phi2p", z1) - e1 * exc.get("phi1p", z1)) - 1j * k * (
            np.einsum("id,idp->ip", nv, exc["a2"]) * e2 - np.einsum("id,idp->ip", nv, exc["a1"]) * e1)
        b = pack(phi, a, De, alpha)

        def afun(x):
            s1, h1, s2, h2 = unpack(x, n)
            ph, aa, al, de = bem_operator_apply(mats, e1, e2, k, nv, {"sig1": s1, "h1": h1, "sig2": s2, "h2": h2})
            return pack(ph, aa, de, al)

        def mfun(x):
            ph, aa, de, al = unpack(x, n)
            sol = _direct_from_parts(d, ph, aa, de, al)
            return pack(sol["sig1"], sol["h1"], sol["sig2"], sol["h2"])

        N = b.size
        A = LinearOperator((N, N), matvec=afun, dtype=complex)
        M = LinearOperator((N, N), matvec=mfun, dtype=complex) if self.precond else None
        restart = self.restart if self.restart is not None else min(N, self.maxit)
        x, flag = gmres(A, b, rtol=self.tol, maxiter=self.maxit, restart=restart, M=M)
        relres = np.linalg.norm(afun(x) - b) / np.linalg.norm(b)
        s1, h1, s2, h2 = unpack(x, n)
        return {"enei": enei, "sig1": s1, "h1": h1, "sig2": s2, "h2": h2, "flag": flag, "relres": relres}


def _direct_from_parts(d, phi, a, De, alpha):
    """The direct solve of bem_ret.py:1694-1810 from already combined right
    sides (the body of bem_ret_iter.py _mfun)."""
    comp, k = d.comp, d.k
    n = comp.n
    nv = comp.nvec
    e1, e2 = d.eps1, d.eps2
    import scipy.linalg as sla
    lsv = lambda lu, bb: sla.lu_solve(lu, bb.reshape(n, -1)).reshape(bb.shape)
    S1 = d.Sigma1
    alpha = alpha - (S1 @ a.reshape(n, -1)).reshape(a.shape) + 1j * k * nv[:, :, None] * (e1 * phi)[:, None, :]
    De = De - S1 @ (e1 * phi) + 1j * k * np.einsum("id,idp->ip", nv, e1 * a)
    Ld = e1 - e2
    sig2 = lsv(d.Slu, De + 1j * k * np.einsum("id,idp->ip", nv, Ld * lsv(d.Dlu, alpha)))
    h2 = lsv(d.Dlu, 1j * k * nv[:, :, None] * (Ld * sig2)[:, None, :] + alpha)
    return {"sig1": lsv(d.G1lu, sig2 + phi), "h1": lsv(d.G1lu, h2 + a),
            "sig2": lsv(d.G2lu, sig2), "h2": lsv(d.G2lu, h2)}


def green_operator(setup, enei):
    """(G1, H1, G2, H2) of the oracle (for tests that apply the operator)."""
    return green_matrices(setup, enei)
