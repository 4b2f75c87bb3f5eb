"""Cluster trees, H-matrices and ACA-compressed Green functions.

``ClusterTree`` (mnpbem/greenfun/clustertree.py) orders the faces by
recursive bisection on the host (one-off set-up, O(N log N)); ``HMatrix``
(mnpbem/greenfun/hmatrix.py) keeps dense leaf blocks and low-rank factors
U V^T of the admissible blocks in device memory, builds the factors with the
partially pivoted ACA on the GPU (one workgroup per admissible block, FP64)
and multiplies with a two-phase GPU matvec; ``ACACompGreenRet``
(mnpbem/greenfun/aca_compgreen_ret.py) compresses the Green matrices of a
CompGreenRet.  As in the reference, ACA entries are entries of the assembled
Green matrix (aca_compgreen_ret.py:164-193 evaluates the dense matrix and
indexes it), so refined near-field entries are included exactly.
"""

import numpy as np
import torch

from . import _native as nat

_DT = {"complex128": (torch.complex128, nat.DTYPE_C128), "complex64": (torch.complex64, nat.DTYPE_C64)}


def _default_fadmiss(eta=2.5):
    return lambda rad1, rad2, dist: eta * min(rad1, rad2) < dist


def make_kaware_fadmiss(k, eta0=2.5):
    """Admissibility eta0 min(r1, r2) < dist, tightened for oscillatory
    kernels: when k min(r1, r2) > 1 the cluster distance must also exceed one
    wavelength 2 pi / |k| (hmatrix.py:1284-1311)."""
    if k is None or not np.isfinite(k) or abs(k) < 1e-12:
        return _default_fadmiss(eta0)
    kabs = abs(complex(k))
    wavelength = 2.0 * np.pi / kabs

    def fadmiss(rad1, rad2, dist):
        rmin = min(rad1, rad2)
        if eta0 * rmin >= dist:
            return False
        if kabs * rmin > 1.0 and dist < wavelength:
            return False
        return True

    return fadmiss


class ClusterTree:
    """Binary cluster tree of face centroids (clustertree.py:7-380).

    A node is split between particles first (faces of the particle of its
    first face against the rest), otherwise at the midpoint of the longest
    side of its bounding box; nodes with more than ``cleaf`` faces (or split
    between particles) are split again.  Node i covers cluster positions
    cind[i, 0] .. cind[i, 1]; ``ind[:, 0]`` maps cluster position -> face,
    ``ind[:, 1]`` face -> cluster position; mid / rad are the bounding-box
    centre and half diagonal; ipart is the particle of the node (-1 when its
    first and last faces belong to different particles)."""

    def __init__(self, pos, cleaf=32, ipart_arr=None):
        self.pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
        self.n = self.pos.shape[0]
        if ipart_arr is None:
            self._ipart_arr = np.zeros(self.n, dtype=np.int64)
        else:
            self._ipart_arr = np.asarray(ipart_arr, dtype=np.int64)
        self.cleaf = cleaf
        self._build(cleaf)

    @staticmethod
    def _sphere(pts):
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        return 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo))

    def _build(self, cleaf):
        n, pos, ip = self.n, self.pos, self._ipart_arr
        perm = np.arange(n, dtype=np.int64)  # cluster position -> face
        m0, r0 = self._sphere(pos)
        son, cind, mid, rad = [[-1, -1]], [[0, n - 1]], [m0], [r0]
        # depth-first: a node's first son is refined completely before its
        # second son, so node numbers follow the reference's recursion
        stack = [(0, 0, n)]
        while stack:
            node, s, e = stack.pop()
            idx = perm[s:e]
            ipn = ip[idx]
            psplit = bool(np.any(ipn != ipn[0]))
            if psplit:
                mask = ipn == ipn[0]
            else:
                pts = pos[idx]
                lo, hi = pts.min(axis=0), pts.max(axis=0)
                ext = hi - lo
                k = int(np.argmax(ext))
                mask = pts[:, k] < lo[k] + 0.5 * ext[k]
            n1 = int(mask.sum())
            if n1 == 0 or n1 == len(idx):  # degenerate cut: peel one face off
                n1 = 1 if n1 == 0 else len(idx) - 1
            else:
                perm[s:e] = np.concatenate([idx[mask], idx[~mask]])
            c1 = len(son)
            son[node] = [c1, c1 + 1]
            for (a, b) in ((s, s + n1), (s + n1, e)):
                mm, rr = self._sphere(pos[perm[a:b]])
                son.append([-1, -1])
                cind.append([a, b - 1])
                mid.append(mm)
                rad.append(rr)
            if (e - s - n1) > cleaf or psplit:
                stack.append((c1 + 1, s + n1, e))
            if n1 > cleaf or psplit:
                stack.append((c1, s, s + n1))
        self.son = np.asarray(son, dtype=np.int64)
        self.cind = np.asarray(cind, dtype=np.int64)
        self.mid = np.asarray(mid, dtype=np.float64)
        self.rad = np.asarray(rad, dtype=np.float64)
        self.ind = np.empty((n, 2), dtype=np.int64)
        self.ind[:, 0] = perm
        self.ind[perm, 1] = np.arange(n)
        a = ip[perm[self.cind[:, 0]]]
        b = ip[perm[self.cind[:, 1]]]
        self.ipart = np.where(a == b, a, -1)

    # ------------------------------------------------------------------
    def leaves(self):
        """Leaf node ids in cluster order."""
        lv = np.nonzero(self.son[:, 0] == -1)[0]
        return lv[np.argsort(self.cind[lv, 0])]

    def cluster_size(self, i):
        return int(self.cind[i, 1] - self.cind[i, 0] + 1)

    def matsize(self, other):
        return (self.n, other.n)

    def matindex(self, other, i1, i2):
        rows = self.ind[self.cind[i1, 0]:self.cind[i1, 1] + 1, 0]
        cols = other.ind[other.cind[i2, 0]:other.cind[i2, 1] + 1, 0]
        return rows, cols, (len(rows), len(cols))

    def part2cluster(self, v):
        return v[self.ind[:, 0]]

    def cluster2part(self, v):
        return v[self.ind[:, 1]]

    def admissibility(self, other=None, fadmiss=None):
        """Block partition {(node1, node2): 1 (low rank) | 2 (dense leaf pair)}
        in the reference's depth-first order (clustertree.py _blocktree)."""
        other = self if other is None else other
        fadmiss = fadmiss if fadmiss is not None else _default_fadmiss()
        out = {}
        stack = [(0, 0)]
        while stack:
            i1, i2 = stack.pop()
            leaf1, leaf2 = self.son[i1, 0] == -1, other.son[i2, 0] == -1
            if leaf1 and leaf2:
                out[(i1, i2)] = 2
                continue
            if self.ipart[i1] != -1 and other.ipart[i2] != -1:
                dist = float(np.linalg.norm(self.mid[i1] - other.mid[i2]))
                if fadmiss(self.rad[i1], other.rad[i2], dist):
                    out[(i1, i2)] = 1
                    continue
            s1 = [i1] if leaf1 else [int(self.son[i1, 0]), int(self.son[i1, 1])]
            s2 = [i2] if leaf2 else [int(other.son[i2, 0]), int(other.son[i2, 1])]
            pairs = [(a, b) for a in s1 for b in s2]
            stack.extend(reversed(pairs))
        return out


class HMatrix:
    """H-matrix over a ClusterTree with device-resident blocks.

    ``dtype`` is the storage type of the blocks ('complex128' or
    'complex64'); products are formed in complex128.  Build with ``aca``
    (from a device matrix, a host array or a host callable fun(row, col)) or
    ``HMatrix.from_func``."""

    def __init__(self, tree=None, htol=1e-6, kmax=100, fadmiss=None, dtype="complex128", device="cuda"):
        if dtype not in _DT:
            raise ValueError("HMatrix: dtype must be 'complex128' or 'complex64'")
        self.tree = tree
        self.htol = htol
        self.kmax = int(max(kmax) if hasattr(kmax, "__iter__") else kmax)
        self.dtype = dtype
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None and torch.cuda.is_available():
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.row1 = self.col1 = self.row2 = self.col2 = np.zeros(0, dtype=np.int64)
        self.ranks = np.zeros(0, dtype=np.int64)
        self._built = False
        if tree is not None:
            self._init(tree, fadmiss)

    def _init(self, tree, fadmiss):
        self.tree = tree
        adm = tree.admissibility(tree, fadmiss=fadmiss)
        keys = list(adm.keys())
        vals = np.array([adm[k] for k in keys], dtype=np.int64)
        kk = np.array(keys, dtype=np.int64).reshape(-1, 2)
        self.row1, self.col1 = kk[vals == 2, 0], kk[vals == 2, 1]
        self.row2, self.col2 = kk[vals == 1, 0], kk[vals == 1, 1]

    # ------------------------------------------------------------------
    def _tables(self):
        t = self.tree
        r1, c1, r2, c2 = self.row1, self.col1, self.row2, self.col2
        nd, nl = len(r1), len(r2)
        blk = np.zeros((nd + nl, 8), dtype=np.int64)
        for (a, rows, cols) in ((0, r1, c1), (nd, r2, c2)):
            blk[a:a + len(rows), 0] = t.cind[rows, 0]
            blk[a:a + len(rows), 1] = t.cind[rows, 1] - t.cind[rows, 0] + 1
            blk[a:a + len(rows), 2] = t.cind[cols, 0]
            blk[a:a + len(rows), 3] = t.cind[cols, 1] - t.cind[cols, 0] + 1
        blk[:nd, 4] = -1
        sizes = blk[:nd, 1] * blk[:nd, 3]
        blk[:nd, 5] = np.concatenate([[0], np.cumsum(sizes)[:-1]]) if nd else []
        self._nval = int(sizes.sum())
        # leaf rows -> blocks (dense blocks of the leaf row, low-rank blocks over it)
        lv = t.leaves()
        starts = t.cind[lv, 0]
        li_d = np.searchsorted(starts, blk[:nd, 0])
        a = np.searchsorted(starts, blk[nd:, 0], "left")
        e = np.searchsorted(starts, blk[nd:, 0] + blk[nd:, 1], "left")
        cnt = e - a
        li_l = (np.repeat(a - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(cnt.sum())
                if nl else np.zeros(0, dtype=np.int64))
        bl_l = np.repeat(np.arange(nd, nd + nl), cnt)
        li = np.concatenate([li_d, li_l]).astype(np.int64)
        bl = np.concatenate([np.arange(nd), bl_l]).astype(np.int64)
        order = np.argsort(li, kind="stable")
        lst = bl[order]
        ptr = np.concatenate([[0], np.cumsum(np.bincount(li, minlength=len(lv)))])
        leafs = np.stack([starts, t.cind[lv, 1] - starts + 1, ptr[:-1], ptr[1:]], axis=1).astype(np.int64)
        self._blk_host = blk
        self._leafs_host = leafs
        self._nd, self._nl = nd, nl
        self._maxchunks = int((leafs[:, 1].max() + 255) // 256) if len(leafs) else 1
        dev = self.device
        self._leafs = torch.as_tensor(leafs).to(dev)
        self._lst = torch.as_tensor(lst.astype(np.int32)).to(dev)
        self._perm = torch.as_tensor(t.ind[:, 0].astype(np.int32)).to(dev)

    def aca(self, fun):
        """Fill the blocks: dense leaf pairs are copied, admissible blocks are
        compressed by ACA on the GPU (hmatrix.py:159-193).  ``fun``: an (n, n)
        complex device tensor (any strides), a host array, or a callable
        fun(row, col) -> values evaluated on the host over all pairs."""
        t = self.tree
        n = t.n
        dev = self.device
        if callable(fun) and not isinstance(fun, (torch.Tensor, np.ndarray)):
            rr, cc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            fun = np.asarray(fun(rr.ravel(), cc.ravel())).reshape(n, n)
        if isinstance(fun, np.ndarray):
            fun = torch.as_tensor(np.ascontiguousarray(fun.astype(np.complex128))).to(dev)
        src = fun if fun.dtype == torch.complex128 else fun.to(torch.complex128)
        if tuple(src.shape) != (n, n) or src.device.type != dev.type:
            raise ValueError("HMatrix.aca: expected an (n, n) matrix on %s" % dev)
        si, sj = src.stride()
        self._tables()
        blk = self._blk_host
        nd, nl = self._nd, self._nl
        tdt, code = _DT[self.dtype]
        st = nat.stream_handle(dev)
        self._val = torch.empty((max(self._nval, 1),), dtype=tdt, device=dev)
        if nd:
            dblk = torch.as_tensor(blk).to(dev)
            bids = torch.arange(nd, dtype=torch.int32, device=dev)
            nat.call("mnpbem_hmat_dense", code, nd, nat.ptr(dblk), nat.ptr(bids), nat.ptr(self._perm),
                     nat.ptr(src), si, sj, nat.ptr(self._val), st)
        ranks = np.zeros(nl, dtype=np.int64)
        if nl:
            if self.kmax > 512:
                raise ValueError("HMatrix.aca: kmax above 512 is not supported")
            m, nn = blk[nd:, 1], blk[nd:, 3]
            if max(m.max(), nn.max()) > 65536:
                raise ValueError("HMatrix.aca: admissible block side above 65536")
            maxr = np.minimum(np.minimum(m, nn), self.kmax)
            ou = np.concatenate([[0], np.cumsum(m * maxr)])
            ov = np.concatenate([[0], np.cumsum(nn * maxr)])
            woff = torch.as_tensor(np.stack([ou[:-1], ov[:-1]], axis=1).astype(np.int64)).to(dev)
            Uw = torch.empty((int(ou[-1]),), dtype=torch.complex128, device=dev)
            Vw = torch.empty((int(ov[-1]),), dtype=torch.complex128, device=dev)
            rk = torch.empty((nl,), dtype=torch.int32, device=dev)
            dblk = torch.as_tensor(blk).to(dev)
            bids = torch.arange(nd, nd + nl, dtype=torch.int32, device=dev)
            nat.call("mnpbem_hmat_aca", nl, nat.ptr(dblk), nat.ptr(bids), nat.ptr(self._perm), nat.ptr(src), si, sj,
                     nat.ptr(Uw), nat.ptr(Vw), nat.ptr(woff), self.kmax, float(self.htol), nat.ptr(rk), st)
            ranks = rk.cpu().numpy().astype(np.int64)
            if np.any(ranks < 0):
                raise RuntimeError("HMatrix.aca: block outside the kernel's capacity")
            blk[nd:, 4] = ranks
            su = np.concatenate([[0], np.cumsum(m * ranks)])
            sv = np.concatenate([[0], np.cumsum(nn * ranks)])
            blk[nd:, 5] = su[:-1]
            blk[nd:, 6] = sv[:-1]
            blk[nd:, 7] = np.concatenate([[0], np.cumsum(ranks)[:-1]])
            self._U = torch.empty((max(int(su[-1]), 1),), dtype=tdt, device=dev)
            self._V = torch.empty((max(int(sv[-1]), 1),), dtype=tdt, device=dev)
            self._blk = torch.as_tensor(blk).to(dev)
            nat.call("mnpbem_hmat_compact", code, nl, nat.ptr(self._blk), nat.ptr(bids), nat.ptr(Uw), nat.ptr(Vw),
                     nat.ptr(woff), nat.ptr(self._U), nat.ptr(self._V), st)
            del Uw, Vw
        else:
            self._U = torch.empty((1,), dtype=tdt, device=dev)
            self._V = torch.empty((1,), dtype=tdt, device=dev)
            self._blk = torch.as_tensor(blk).to(dev)
        self.ranks = ranks
        self._rsum = int(ranks.sum())
        self._T = torch.empty((max(self._rsum, 1) * 16,), dtype=torch.complex128, device=dev)
        self._built = True
        return self

    @staticmethod
    def from_func(tree, fun, htol=1e-6, kmax=100, fadmiss=None, dtype="complex128", device="cuda"):
        return HMatrix(tree, htol=htol, kmax=kmax, fadmiss=fadmiss, dtype=dtype, device=device).aca(fun)

    # ------------------------------------------------------------------
    def _check(self):
        if not self._built:
            raise RuntimeError("HMatrix: blocks not filled (call aca first)")

    def matvec_device(self, X, ldx, Y, ldy, ncol, accumulate=False, stream=None):
        """Y (+)= M X for complex128 device buffers in face order, (n, ldx) and
        (n, ldy) row-major, ncol <= 16 columns."""
        self._check()
        st = nat.stream_handle(self.device) if stream is None else stream
        nat.call("mnpbem_hmat_matvec", _DT[self.dtype][1], len(self._blk_host), len(self._leafs_host),
                 self._maxchunks, nat.ptr(self._blk), nat.ptr(self._leafs), nat.ptr(self._lst), nat.ptr(self._perm),
                 nat.ptr(self._val), nat.ptr(self._U), nat.ptr(self._V), ncol, X, ldx, Y, ldy,
                 1 if accumulate else 0, nat.ptr(self._T), st)

    def mtimes_vec(self, v):
        """M v for v (n,) or (n, k), numpy or torch; the result has the type
        and device of v (hmatrix.py:450-541)."""
        self._check()
        as_np = isinstance(v, np.ndarray)
        x = torch.as_tensor(v) if as_np else v
        one = x.dim() == 1
        x2 = x.reshape(x.shape[0], -1).to(self.device, torch.complex128).contiguous()
        k = x2.shape[1]
        y = torch.empty_like(x2)
        for c0 in range(0, k, 16):
            c1 = min(k, c0 + 16)
            xs = x2[:, c0:c1].contiguous()
            ys = torch.empty_like(xs)
            self.matvec_device(nat.ptr(xs), c1 - c0, nat.ptr(ys), c1 - c0, c1 - c0)
            y[:, c0:c1] = ys
        y = y[:, 0] if one else y
        if as_np:
            return y.cpu().numpy()
        return y.to(x.device)

    def __matmul__(self, other):
        if isinstance(other, HMatrix):
            raise NotImplementedError("HMatrix @ HMatrix is not provided")
        return self.mtimes_vec(other)

    def full(self, xp=None):
        """Dense (n, n) complex128 matrix (hmatrix.py:338-448): a device tensor,
        or a numpy array when xp is numpy."""
        self._check()
        n = self.tree.n
        out = torch.zeros((n, n), dtype=torch.complex128, device=self.device)
        nat.call("mnpbem_hmat_full", _DT[self.dtype][1], len(self._blk_host), nat.ptr(self._blk),
                 nat.ptr(self._perm), nat.ptr(self._val), nat.ptr(self._U), nat.ptr(self._V), nat.ptr(out), n,
                 nat.stream_handle(self.device))
        return out.cpu().numpy() if xp is np else out

    def compression(self):
        """Stored elements / n^2 (hmatrix.py:739-756; a rank-0 block counts as
        rank 1, as the reference stores zero factors for it)."""
        blk = self._blk_host
        nd = self._nd
        dense = int((blk[:nd, 1] * blk[:nd, 3]).sum())
        r = np.maximum(blk[nd:, 4], 1)
        low = int(((blk[nd:, 1] + blk[nd:, 3]) * r).sum())
        return (dense + low) / float(self.tree.n ** 2)

    def stat(self):
        """Block counts, ranks and memory."""
        blk = self._blk_host
        nd = self._nd
        el = np.dtype(np.complex128 if self.dtype == "complex128" else np.complex64).itemsize
        return {"n_dense": int(nd), "n_lowrank": int(len(blk) - nd), "rank_max": int(self.ranks.max()) if len(self.ranks) else 0,
                "rank_mean": float(self.ranks.mean()) if len(self.ranks) else 0.0,
                "compression": self.compression(),
                "bytes": int((self._nval + (blk[nd:, 1] * blk[nd:, 4]).sum() + (blk[nd:, 3] * blk[nd:, 4]).sum()) * el)}

    def bytes_per_matvec(self):
        """Block bytes one matvec reads (dense values plus both factors)."""
        return self.stat()["bytes"]


class ACACompGreenRet:
    """Green functions of a ComParticle compressed to H-matrices
    (aca_compgreen_ret.py:96-208).  ``eval(i, j, key, enei)`` returns an
    HMatrix of the dense CompGreenRet.eval matrix (or the int 0 for regions
    that no particle connects); the default admissibility is the
    wavenumber-aware rule of make_kaware_fadmiss(2 pi / enei, eta)."""

    def __init__(self, p, htol=1e-6, kmax=100, cleaf=32, fadmiss=None, eta=2.5, dtype="complex128", green=None,
                 **options):
        from .greenfun import CompGreenRet
        self.p = p
        self.g = green if green is not None else CompGreenRet(p, p, **options)
        self.tree = ClusterTree(p.pos, cleaf=cleaf, ipart_arr=self._ipart(p))
        self.eta = eta
        self._user_fadmiss = fadmiss
        self.htol, self.kmax, self.cleaf, self.dtype = htol, kmax, cleaf, dtype
        self._cache = {}

    @staticmethod
    def _ipart(p):
        if len(getattr(p, "p", [])) <= 1:
            return None
        return np.asarray(p.surface, dtype=np.int64)

    def fadmiss_for(self, enei):
        if self._user_fadmiss is not None:
            return self._user_fadmiss
        return make_kaware_fadmiss(2.0 * np.pi / enei, eta0=self.eta)

    def compress(self, M, enei):
        """HMatrix of an assembled (n, n) device matrix at wavelength enei."""
        h = HMatrix(self.tree, htol=self.htol, kmax=self.kmax, fadmiss=self.fadmiss_for(enei), dtype=self.dtype,
                    device=self.g.device)
        return h.aca(M)

    def eval(self, i, j, key, enei):
        ck = (i, j, key, float(enei))
        if ck in self._cache:
            return self._cache[ck]
        M = self.g.eval(i, j, key, enei)
        if isinstance(M, int):
            return 0
        h = self.compress(M, enei)
        self._cache[ck] = h
        return h

    def eval_full(self, i, j, key, enei):
        h = self.eval(i, j, key, enei)
        return h if isinstance(h, int) else h.full()

    def compression(self, i=0, j=0, key="G", enei=600.0):
        return self.eval(i, j, key, enei).compression()

    def clear_cache(self):
        self._cache = {}
