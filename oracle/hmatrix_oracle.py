"""FP64 CPU oracle of the cluster tree, H-matrix and ACA path.

TEST INFRASTRUCTURE ONLY (imported by ``tests/`` as the checker).

Restates mnpbem/greenfun/clustertree.py:7-380 (recursive bisection, block
partition), mnpbem/greenfun/hmatrix.py:159-336 (HMatrix.aca, _aca_block),
:338-541 (full, mtimes_vec), :672-737 (truncate, _truncate_block),
:739-756 (compression) and :1284-1311 (make_kaware_fadmiss) in plain numpy.
The tree is built recursively, node by node, as in the reference; ACA runs
on a host callable fun(row, col).

Pinning: tests/test_hmatrix_oracle.py checks the tree invariants the
reference's own tests assert (mnpbem/tests/test_hmatrix.py: leaf sizes,
bijective index maps, bounding spheres, leaf-pair coverage of the matrix),
ACA accuracy on the 1/(4 pi r) kernel against the dense matrix, exact
recovery of a matrix of known rank, truncation against the SVD of the
block, and the k-aware admissibility against its definition.
"""

import numpy as np


def _bounding_sphere(pts):
    """clustertree.py _sph_boundary: centre and half diagonal of the box."""
    lo = pts.min(axis=0)
    hi = pts.max(axis=0)
    return 0.5 * (lo + hi), 0.5 * np.linalg.norm(hi - lo)


class OracleClusterTree:
    """clustertree.py:7-380.  Nodes are dicts appended as the recursion
    visits them; son1 is refined before son2."""

    def __init__(self, pos, cleaf=32, ipart_arr=None):
        self.pos = np.asarray(pos, dtype=np.float64)
        self.n = len(self.pos)
        self.ipart_arr = (np.zeros(self.n, dtype=np.int64) if ipart_arr is None
                          else np.asarray(ipart_arr, dtype=np.int64))
        mid, rad = _bounding_sphere(self.pos)
        self.nodes = [dict(cs=0, ce=self.n - 1, mid=mid, rad=rad, son=(-1, -1), faces=None)]
        self._refine(0, np.arange(self.n), np.arange(self.n), cleaf)
        self.son = np.array([nd["son"] for nd in self.nodes], dtype=np.int64)
        self.cind = np.array([[nd["cs"], nd["ce"]] for nd in self.nodes], dtype=np.int64)
        self.mid = np.array([nd["mid"] for nd in self.nodes])
        self.rad = np.array([nd["rad"] for nd in self.nodes])
        c2p = np.empty(self.n, dtype=np.int64)
        for nd in self.nodes:
            if nd["faces"] is not None:
                c2p[nd["cs"]:nd["ce"] + 1] = nd["faces"]
        self.ind = np.stack([c2p, np.argsort(c2p)], axis=1)
        ip = self.ipart_arr
        self.ipart = np.array([ip[c2p[nd["cs"]]] if ip[c2p[nd["cs"]]] == ip[c2p[nd["ce"]]] else -1
                               for nd in self.nodes], dtype=np.int64)

    def _refine(self, ic, ind, cind, cleaf):
        ip = self.ipart_arr[ind]
        psplit = len(np.unique(ip)) > 1
        if psplit:
            sel = ip == ip[0]
            ind1, ind2 = ind[sel], ind[~sel]
        else:
            p = self.pos[ind]
            ext = p.max(axis=0) - p.min(axis=0)
            k = int(np.argmax(ext))
            cut = p.min(axis=0)[k] + 0.5 * ext[k]
            sel = self.pos[ind, k] < cut
            ind1, ind2 = ind[sel], ind[~sel]
            if len(ind1) == 0:
                ind1, ind2 = ind[:1], ind[1:]
            elif len(ind2) == 0:
                ind1, ind2 = ind[:-1], ind[-1:]
        c1, c2 = cind[:len(ind1)], cind[len(ind1):]
        base = len(self.nodes)
        self.nodes[ic]["son"] = (base, base + 1)
        for faces, cc in ((ind1, c1), (ind2, c2)):
            mid, rad = _bounding_sphere(self.pos[faces])
            self.nodes.append(dict(cs=int(cc[0]), ce=int(cc[-1]), mid=mid, rad=rad, son=(-1, -1), faces=None))
        for off, faces, cc in ((0, ind1, c1), (1, ind2, c2)):
            if len(faces) > cleaf or psplit:
                self._refine(base + off, faces, cc, cleaf)
            else:
                self.nodes[base + off]["faces"] = faces

    def is_leaf(self, i):
        return self.son[i, 0] == -1

    def admissibility(self, fadmiss=None):
        """clustertree.py admissibility/_blocktree: {(i1, i2): 1 | 2}."""
        fadmiss = fadmiss or (lambda r1, r2, d: 2.5 * min(r1, r2) < d)
        out = {}

        def visit(i1, i2):
            if self.is_leaf(i1) and self.is_leaf(i2):
                out[(i1, i2)] = 2
                return
            if self.ipart[i1] != -1 and self.ipart[i2] != -1 and \
                    fadmiss(self.rad[i1], self.rad[i2], np.linalg.norm(self.mid[i1] - self.mid[i2])):
                out[(i1, i2)] = 1
                return
            for a in ([i1] if self.is_leaf(i1) else list(self.son[i1])):
                for b in ([i2] if self.is_leaf(i2) else list(self.son[i2])):
                    visit(int(a), int(b))

        visit(0, 0)
        return out


def kaware_fadmiss(k, eta0=2.5):
    """hmatrix.py:1284-1311."""
    if k is None or not np.isfinite(k) or abs(k) < 1e-12:
        return lambda r1, r2, d: eta0 * min(r1, r2) < d
    lam = 2.0 * np.pi / abs(k)

    def f(r1, r2, d):
        rmin = min(r1, r2)
        return not (eta0 * rmin >= d or (abs(k) * rmin > 1.0 and d < lam))

    return f


def aca_block(fun, rows, cols, htol, kmax):
    """hmatrix.py:195-336: partially pivoted ACA, Frobenius-norm stop."""
    m, n = len(rows), len(cols)
    U, V = [], []
    used_r = np.zeros(m, bool)
    used_c = np.zeros(n, bool)
    p = 0
    fro2 = 0.0
    for _ in range(min(m, n, kmax)):
        row = np.asarray(fun(np.full(n, rows[p]), cols), dtype=complex)
        for u, v in zip(U, V):
            row = row - u[p] * v
        a = np.abs(row)
        a[used_c] = 0.0
        q = int(np.argmax(a))
        piv = row[q]
        if abs(piv) < 1e-15:
            break
        col = np.asarray(fun(rows, np.full(m, cols[q])), dtype=complex)
        for u, v in zip(U, V):
            col = col - u * v[q]
        u_new, v_new = col / piv, row
        cross = sum(np.vdot(u, u_new) * np.vdot(v, v_new) for u, v in zip(U, V))
        U.append(u_new)
        V.append(v_new)
        used_r[p] = True
        used_c[q] = True
        nt = np.vdot(u_new, u_new).real * np.vdot(v_new, v_new).real
        fro2 += nt + 2.0 * np.real(cross)
        if fro2 > 0 and np.sqrt(nt) < htol * np.sqrt(abs(fro2)):
            break
        a = np.abs(u_new)
        a[used_r] = 0.0
        p = int(np.argmax(a))
    if not U:
        return np.zeros((m, 1), complex), np.zeros((n, 1), complex)
    return np.array(U).T, np.array(V).T


def truncate_block(lhs, rhs, htol):
    """hmatrix.py:688-737: QR of both factors, SVD of R1 R2^T, keep the
    leading singular values holding a (1 - htol) share of their sum."""
    if np.linalg.norm(lhs) < np.finfo(float).eps or np.linalg.norm(rhs) < np.finfo(float).eps:
        return lhs, rhs
    q1, r1 = np.linalg.qr(lhs)
    q2, r2 = np.linalg.qr(rhs)
    u, s, vt = np.linalg.svd(r1 @ r2.T, full_matrices=False)
    tot = s.sum()
    if tot < np.finfo(float).eps:
        return lhs[:, :1] * 0, rhs[:, :1] * 0
    k = max(1, int(np.count_nonzero(np.cumsum(s) < (1.0 - htol) * tot)))
    return q1 @ (u[:, :k] * s[:k]), q2 @ vt[:k, :].T


class OracleHMatrix:
    """hmatrix.py HMatrix: dense leaf pairs and ACA low-rank blocks."""

    def __init__(self, tree, htol=1e-6, kmax=100, fadmiss=None):
        self.tree, self.htol, self.kmax = tree, htol, kmax
        adm = tree.admissibility(fadmiss)
        self.dense = [k for k, v in adm.items() if v == 2]
        self.lowrank = [k for k, v in adm.items() if v == 1]
        self.val, self.lhs, self.rhs = [], [], []

    def _rc(self, i):
        cs, ce = self.tree.cind[i]
        return self.tree.ind[cs:ce + 1, 0]

    def aca(self, fun):
        for (i1, i2) in self.dense:
            r, c = self._rc(i1), self._rc(i2)
            rr, cc = np.meshgrid(r, c, indexing="ij")
            self.val.append(np.asarray(fun(rr.ravel(), cc.ravel()), dtype=complex).reshape(rr.shape))
        for (i1, i2) in self.lowrank:
            u, v = aca_block(fun, self._rc(i1), self._rc(i2), self.htol, self.kmax)
            self.lhs.append(u)
            self.rhs.append(v)
        return self

    def full(self):
        n = self.tree.n
        out = np.zeros((n, n), complex)
        for (i1, i2), b in zip(self.dense, self.val):
            out[np.ix_(self._rc(i1), self._rc(i2))] = b
        for (i1, i2), u, v in zip(self.lowrank, self.lhs, self.rhs):
            out[np.ix_(self._rc(i1), self._rc(i2))] = u @ v.T
        return out

    def mtimes_vec(self, x):
        x = np.asarray(x, dtype=complex)
        y = np.zeros_like(x)
        for (i1, i2), b in zip(self.dense, self.val):
            y[self._rc(i1)] += b @ x[self._rc(i2)]
        for (i1, i2), u, v in zip(self.lowrank, self.lhs, self.rhs):
            y[self._rc(i1)] += u @ (v.T @ x[self._rc(i2)])
        return y

    def compression(self):
        el = sum(b.size for b in self.val) + sum(u.size + v.size for u, v in zip(self.lhs, self.rhs))
        return el / float(self.tree.n ** 2)

    def ranks(self):
        return np.array([0 if not np.any(u) else u.shape[1] for u in self.lhs], dtype=np.int64)

    def truncate(self, htol=None):
        htol = self.htol if htol is None else htol
        for i in range(len(self.lhs)):
            self.lhs[i], self.rhs[i] = truncate_block(self.lhs[i], self.rhs[i], htol)
        return self
