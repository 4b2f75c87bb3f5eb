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


def pack_tables(blk, nd, n, leafs, lst):
    """Piece and item tables of the packed product layouts for a block table
    ``blk`` (int64, 8 per block: r0, m, c0, n, rank, off_a, off_b, off_t;
    the first ``nd`` dense) and leaf table ``leafs`` / ``lst``:

    * column panels: the low-rank blocks of rank > 0 that share a column
      cluster, their V factors side by side and transposed,
      VS[vs + j K + o] = V_b[j][r] for output o = (b, r) of the panel, whose
      t value goes to workspace row tix[og + o] = n + off_t + r.  Pieces
      (one per block) {dst, K, n, off_b, rank, tdst, trow, 0}; items (one per
      64 outputs of a panel) {vs, K, o0, n, c0, og}.
    * row panels: per leaf row cluster and 64-row group (mi rows), the rows
      of every block over it side by side, column-major,
      P[p + q mi + i]; column q multiplies workspace row gidx[g + q] (X row
      c0 + j of a dense block, t row n + off_t + r of a low-rank one).
      Pieces (one per group and block) {dst, mi, ncols, src, ld, dense,
      gdst, gstart}; items {p, mi, K, g, r0, 0}.

    Returns (cdesc, citems, rdesc, ritems, (len VS, len tix, len P, len gidx))."""
    i64 = np.int64
    lr = np.nonzero(blk[nd:, 4] > 0)[0] + nd
    cdesc = np.zeros((0, 8), i64)
    citems = np.zeros((0, 6), i64)
    nvs = ntix = 0
    if len(lr):
        lr = lr[np.argsort(blk[lr, 2] * (n + 1) + blk[lr, 3], kind="stable")]
        key = blk[lr, 2] * (n + 1) + blk[lr, 3]
        k = blk[lr, 4]
        newg = np.r_[True, key[1:] != key[:-1]]
        gid = np.cumsum(newg) - 1
        gs = np.nonzero(newg)[0]
        Kg = np.add.reduceat(k, gs)
        nG, c0G = blk[lr[gs], 3], blk[lr[gs], 2]
        vsG = np.r_[0, np.cumsum(nG * Kg)[:-1]].astype(i64)
        ogG = np.r_[0, np.cumsum(Kg)[:-1]].astype(i64)
        kx = np.cumsum(k) - k
        ooff = kx - kx[gs][gid]
        cdesc = np.zeros((len(lr), 8), i64)
        cdesc[:, 0] = vsG[gid] + ooff
        cdesc[:, 1] = Kg[gid]
        cdesc[:, 2] = blk[lr, 3]
        cdesc[:, 3] = blk[lr, 6]
        cdesc[:, 4] = k
        cdesc[:, 5] = ogG[gid] + ooff
        cdesc[:, 6] = n + blk[lr, 7]
        nsl = (Kg + 63) // 64
        gi = np.repeat(np.arange(len(gs)), nsl)
        sl = np.arange(int(nsl.sum())) - np.repeat(np.cumsum(nsl) - nsl, nsl)
        citems = np.stack([vsG[gi], Kg[gi], 64 * sl, nG[gi], c0G[gi], ogG[gi]], axis=1).astype(i64)
        # longest panels first: the workgroups start in item order, so the
        # long items do not form the tail of the launch
        citems = citems[np.argsort(-citems[:, 3], kind="stable")]
        nvs, ntix = int((nG * Kg).sum()), int(Kg.sum())
    ngr = (leafs[:, 1] + 63) // 64
    li = np.repeat(np.arange(len(leafs)), ngr)
    gg = np.arange(int(ngr.sum())) - np.repeat(np.cumsum(ngr) - ngr, ngr)
    ir0 = leafs[li, 0] + 64 * gg
    imi = np.minimum(64, leafs[li, 1] - 64 * gg)
    cnt = leafs[li, 3] - leafs[li, 2]
    pi = np.repeat(np.arange(len(li)), cnt)
    pe = np.repeat(leafs[li, 2], cnt) + (np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    pb = np.asarray(lst)[pe].astype(i64)
    keep = blk[pb, 4] != 0
    pi, pb = pi[keep], pb[keep]
    rk = blk[pb, 4]
    dense = rk < 0
    ncols = np.where(dense, blk[pb, 3], rk)
    K = np.bincount(pi, weights=ncols, minlength=len(li)).astype(i64)
    gofs = np.r_[0, np.cumsum(K)[:-1]].astype(i64)
    pcum = np.cumsum(ncols) - ncols
    pbase = np.r_[0, np.cumsum(imi * K)[:-1]].astype(i64)
    rdesc = np.zeros((len(pb), 8), i64)
    rdesc[:, 0] = pbase[pi] + (pcum - gofs[pi]) * imi[pi]
    rdesc[:, 1] = imi[pi]
    rdesc[:, 2] = ncols
    rdesc[:, 3] = blk[pb, 5] + (ir0[pi] - blk[pb, 0])
    rdesc[:, 4] = blk[pb, 1]
    rdesc[:, 5] = dense
    rdesc[:, 6] = pcum
    rdesc[:, 7] = np.where(dense, blk[pb, 2], n + blk[pb, 7])
    ritems = np.stack([pbase, imi, K, gofs, ir0, np.zeros_like(K)], axis=1).astype(i64)
    ritems = ritems[np.argsort(-(imi * K), kind="stable")]
    return cdesc, citems, rdesc, ritems, (nvs, ntix, int((imi * K).sum()), int(K.sum()))


class HMatrix:
    """H-matrix over a ClusterTree with device-resident blocks.

    ``dtype`` is the storage type of the blocks ('complex128' or
    'complex64'); products are formed in complex128.  Build with ``aca``
    (from a device matrix, a host array or a host callable fun(row, col)) or
    ``HMatrix.from_func``."""

    def __init__(self, tree=None, htol=1e-6, kmax=100, fadmiss=None, dtype="complex128", device="cuda",
                 partition=None):
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
        self._lu = None
        self._pk = None
        if tree is not None and partition is not None:
            # the (dense, low-rank) block lists of another H-matrix over the
            # same tree and admissibility
            self.row1, self.col1, self.row2, self.col2 = partition
        elif tree is not None:
            self._init(tree, fadmiss)

    @property
    def partition(self):
        """(row1, col1, row2, col2): the dense and low-rank cluster pairs."""
        return self.row1, self.col1, self.row2, self.col2

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
        self._lst_host = lst
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

        def dense(code, nb, dblk, bids, val, st):
            nat.call("mnpbem_hmat_dense", code, nb, dblk, bids, nat.ptr(self._perm), nat.ptr(src), si, sj, val, st)

        def aca(nb, dblk, bids, Uw, Vw, woff, rk, st):
            nat.call("mnpbem_hmat_aca", nb, dblk, bids, nat.ptr(self._perm), nat.ptr(src), si, sj, Uw, Vw, woff,
                     self.kmax, float(self.htol), rk, st)

        return self._fill(dense, aca)

    def aca_green(self, green, side, key, kmat):
        """Fill the blocks with entries generated in place from the geometry
        and refinement tables of CompGreenRet ``green`` (side 0 inside, 1
        outside; key 'G', 'F', 'H1' or 'H2'; kmat: the (nmat, 2) float64
        device wavenumbers of one wavelength, a row of green.wavenumbers):
        the dense leaf blocks and the ACA rows and columns of the admissible
        blocks are computed by the kernels, so no (n, n) matrix is assembled
        (the entry functions of mex/acagreen/acagreen.cpp:79-127, with the
        refined near-field entries of compgreen_ret.py:507-553)."""
        kinds = {"G": (0, 0.0), "F": (1, 0.0), "H1": (1, 1.0), "H2": (1, -1.0)}
        if key not in kinds:
            raise ValueError("HMatrix.aca_green: key must be one of %s" % sorted(kinds))
        if green.n != self.tree.n:
            raise ValueError("HMatrix.aca_green: the tree and the Green function have different face counts")
        kind, hs = kinds[key]
        kmat = kmat.contiguous()
        geo = (nat.ptr(green.pos), nat.ptr(green.nvec), nat.ptr(green.area), nat.ptr(green.mat[side]),
               nat.ptr(kmat), green.order, nat.ptr(green._rowptr()), nat.ptr(green.pcol), nat.ptr(green.g),
               nat.ptr(green.f))

        def dense(code, nb, dblk, bids, val, st):
            nat.call("mnpbem_hmat_dense_green", code, nb, dblk, bids, nat.ptr(self._perm), kind, hs, *geo, val, st)

        def aca(nb, dblk, bids, Uw, Vw, woff, rk, st):
            nat.call("mnpbem_hmat_aca_green", nb, dblk, bids, nat.ptr(self._perm), kind, hs, *geo, Uw, Vw, woff,
                     self.kmax, float(self.htol), rk, st)

        return self._fill(dense, aca)

    def _fill(self, dense, aca):
        """Dense leaf blocks through ``dense`` and ACA of the admissible
        blocks through ``aca`` (launchers of one entry source), then the
        packed factors."""
        dev = self.device
        self._tables()
        blk = self._blk_host
        nd, nl = self._nd, self._nl
        tdt, code = _DT[self.dtype]
        st = nat.stream_handle(dev)
        self._val = torch.empty((max(self._nval, 1),), dtype=tdt, device=dev)
        if nd:
            dblk = torch.as_tensor(blk).to(dev)
            bids = torch.arange(nd, dtype=torch.int32, device=dev)
            dense(code, nd, nat.ptr(dblk), nat.ptr(bids), nat.ptr(self._val), st)
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
            aca(nl, nat.ptr(dblk), nat.ptr(bids), nat.ptr(Uw), nat.ptr(Vw), nat.ptr(woff), nat.ptr(rk), st)
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
        self._lu = None
        self._pk = None
        return self

    @staticmethod
    def from_func(tree, fun, htol=1e-6, kmax=100, fadmiss=None, dtype="complex128", device="cuda"):
        return HMatrix(tree, htol=htol, kmax=kmax, fadmiss=fadmiss, dtype=dtype, device=device).aca(fun)

    # ------------------------------------------------------------------
    def _check(self, blocks=True):
        if not self._built:
            raise RuntimeError("HMatrix: blocks not filled (call aca first)")
        if blocks and self._val is None:
            raise RuntimeError("HMatrix: block storage released (release_blocks); only products remain")

    def release_blocks(self):
        """Build the packed product layouts and free the block storage (dense
        leaf values and low-rank factors), for an H-matrix used only as an
        operator: its device memory halves, products (matvec_device,
        mtimes_vec) and the host tables (stat, compression) remain, and
        every method that reads or changes the blocks raises."""
        self._check()
        if self._pk is None:
            self._pack()
        self._val = self._U = self._V = None
        return self

    def _pack(self):
        """Packed product layouts, built on the first product after the
        blocks change (see pack_tables)."""
        n = self.tree.n
        dev = self.device
        tdt, code = _DT[self.dtype]
        cdesc, citems, rdesc, ritems, (nvs, ntix, nP, ng) = pack_tables(self._blk_host, self._nd, n,
                                                                       self._leafs_host, self._lst_host)
        if max(n + self._rsum, ng, ntix) >= 2 ** 31:
            raise ValueError("HMatrix: product tables above 2^31 entries")
        pk = {"VS": torch.empty((max(nvs, 1),), dtype=tdt, device=dev),
              "tix": torch.empty((max(ntix, 1),), dtype=torch.int32, device=dev),
              "P": torch.empty((max(nP, 1),), dtype=tdt, device=dev),
              "gidx": torch.empty((max(ng, 1),), dtype=torch.int32, device=dev),
              "citems": torch.as_tensor(citems).to(dev), "ritems": torch.as_tensor(ritems).to(dev),
              "nc": len(citems), "nr": len(ritems),
              "R": torch.empty(((n + self._rsum) * 16,), dtype=torch.complex128, device=dev)}
        cd, rd = torch.as_tensor(cdesc).to(dev), torch.as_tensor(rdesc).to(dev)
        nat.call("mnpbem_hmat_pack", code, len(cdesc), nat.ptr(cd), len(rdesc), nat.ptr(rd), nat.ptr(self._val),
                 nat.ptr(self._U), nat.ptr(self._V), nat.ptr(pk["VS"]), nat.ptr(pk["tix"]), nat.ptr(pk["P"]),
                 nat.ptr(pk["gidx"]), nat.stream_handle(dev))
        pk["_desc"] = (cd, rd)  # kept with the layouts: the pack launches read them asynchronously
        self._pk = pk
        return pk

    def matvec_device(self, X, ldx, Y, ldy, ncol, accumulate=False, stream=None):
        """Y (+)= M X for complex128 device buffers in face order, (n, ldx) and
        (n, ldy) row-major, ncol <= 16 columns (on the packed layouts)."""
        self._check(blocks=self._pk is None)
        st = nat.stream_handle(self.device) if stream is None else stream
        pk = self._pk if self._pk is not None else self._pack()
        nat.call("mnpbem_hmat_matvec_packed", _DT[self.dtype][1], self.tree.n, pk["nc"], nat.ptr(pk["citems"]),
                 nat.ptr(pk["VS"]), nat.ptr(pk["tix"]), pk["nr"], nat.ptr(pk["ritems"]), nat.ptr(pk["P"]),
                 nat.ptr(pk["gidx"]), nat.ptr(self._perm), ncol, X, ldx, Y, ldy, 1 if accumulate else 0,
                 nat.ptr(pk["R"]), st)

    def matvec_blocks_device(self, X, ldx, Y, ldy, ncol, accumulate=False, stream=None):
        """As matvec_device, on the block storage itself (one workgroup per
        leaf row cluster walking its blocks)."""
        self._check()
        st = nat.stream_handle(self.device) if stream is None else stream
        if getattr(self, "_XW", None) is None or self._XW.numel() < self.tree.n * 16:
            self._XW = torch.empty((self.tree.n * 16,), dtype=torch.complex128, device=self.device)
        nat.call("mnpbem_hmat_matvec", _DT[self.dtype][1], self.tree.n, len(self._blk_host), len(self._leafs_host),
                 self._maxchunks, nat.ptr(self._blk), nat.ptr(self._leafs), nat.ptr(self._lst), nat.ptr(self._perm),
                 nat.ptr(self._val), nat.ptr(self._U), nat.ptr(self._V), ncol, X, ldx, Y, ldy,
                 1 if accumulate else 0, nat.ptr(self._T), nat.ptr(self._XW), st)

    def mtimes_vec(self, v):
        """M v for v (n,) or (n, k), numpy or torch; the result has the type
        and device of v (hmatrix.py:450-541)."""
        self._check(blocks=self._pk is None)
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
            return self._mtimes_hmat(other)
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

    # ------------------------------------------------------- block inspection
    def _host_blocks(self, which):
        self._check()
        blk = self._blk_host
        nd = self._nd
        if which == "val":
            h = self._val.cpu().numpy().astype(np.complex128)
            return [h[b[5]:b[5] + b[1] * b[3]].reshape(b[3], b[1]).T for b in blk[:nd]]
        buf = (self._U if which == "lhs" else self._V).cpu().numpy().astype(np.complex128)
        col, off = (1, 5) if which == "lhs" else (3, 6)
        return [buf[b[off]:b[off] + b[col] * max(b[4], 0)].reshape(max(b[4], 0), b[col]).T for b in blk[nd:]]

    @property
    def val(self):
        """Dense blocks as host arrays (row1/col1 order, rows and columns in
        cluster order), as HMatrix.val of hmatrix.py."""
        return self._host_blocks("val")

    @property
    def lhs(self):
        """Left factors U (m x rank) of the low-rank blocks, host arrays."""
        return self._host_blocks("lhs")

    @property
    def rhs(self):
        """Right factors V (n x rank), block = U V^T, host arrays."""
        return self._host_blocks("rhs")

    # ------------------------------------------------------------------ algebra
    def _set_lowrank(self, blk, U, V):
        """Install a block table whose low-rank ranks (column 4) changed:
        recompute the factor offsets and the matvec workspace."""
        nd = self._nd
        m, nn, r = blk[nd:, 1], blk[nd:, 3], np.maximum(blk[nd:, 4], 0)
        blk[nd:, 5] = np.concatenate([[0], np.cumsum(m * r)[:-1]]) if len(r) else []
        blk[nd:, 6] = np.concatenate([[0], np.cumsum(nn * r)[:-1]]) if len(r) else []
        blk[nd:, 7] = np.concatenate([[0], np.cumsum(r)[:-1]]) if len(r) else []
        self._blk_host = blk
        self._blk = torch.as_tensor(blk).to(self.device)
        self._U, self._V = U, V
        self.ranks = blk[nd:, 4].copy()
        self._rsum = int(r.sum())
        self._T = torch.empty((max(self._rsum, 1) * 16,), dtype=torch.complex128, device=self.device)
        self._pk = None

    def _factor_bufs(self, blk):
        nd = self._nd
        r = np.maximum(blk[nd:, 4], 0)
        tdt = _DT[self.dtype][0]
        nu = int((blk[nd:, 1] * r).sum())
        nv = int((blk[nd:, 3] * r).sum())
        return (torch.empty((max(nu, 1),), dtype=tdt, device=self.device),
                torch.empty((max(nv, 1),), dtype=tdt, device=self.device))

    def _copy(self):
        """Deep copy: same tree and tables, own device blocks."""
        self._check()
        c = HMatrix.__new__(HMatrix)
        c.__dict__.update(self.__dict__)
        c._val, c._U, c._V = self._val.clone(), self._U.clone(), self._V.clone()
        c._blk_host = self._blk_host.copy()
        c._blk = self._blk.clone()
        c.ranks = self.ranks.copy()
        c._T = torch.empty_like(self._T)
        c._lu = None
        c._pk = None
        return c

    def truncate(self, htol=None):
        """Recompress every low-rank block in place (hmatrix.py:672-737): QR of
        both factors, SVD of R1 R2^T, keep the leading singular values that
        hold a (1 - htol) share of their sum (at least one).  Runs on the GPU
        in FP64 for ranks up to 128."""
        self._check()
        htol = self.htol if htol is None else float(htol)
        blk = self._blk_host
        nd = self._nd
        lr = np.nonzero(blk[nd:, 4] > 0)[0] + nd
        if len(lr):
            k = blk[lr, 4]
            kst = int(k.max())
            if kst > 128:
                raise ValueError("HMatrix.truncate: ranks above 128 are not supported (largest rank %d)" % kst)
            dev = self.device
            st = nat.stream_handle(dev)
            code = _DT[self.dtype][1]
            nb = len(lr)
            sizes = np.stack([blk[lr, 1] * k, blk[lr, 3] * k], axis=1).ravel()
            qoff = torch.as_tensor(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)).to(dev)
            Q = torch.empty((int(sizes.sum()),), dtype=torch.complex128, device=dev)
            RT = torch.empty((2 * nb * kst * kst,), dtype=torch.complex128, device=dev)
            XZ = torch.empty((2 * nb * kst * kst,), dtype=torch.complex128, device=dev) if kst > 64 else None
            bids = torch.as_tensor(lr.astype(np.int32)).to(dev)
            rk = torch.empty((nb,), dtype=torch.int32, device=dev)
            nat.call("mnpbem_hmat_truncate", code, nb, nat.ptr(self._blk), nat.ptr(bids), nat.ptr(self._U),
                     nat.ptr(self._V), nat.ptr(qoff), kst, htol, nat.ptr(Q), nat.ptr(RT),
                     nat.ptr(XZ) if XZ is not None else None, nat.ptr(rk), st)
            r = rk.cpu().numpy().astype(np.int64)
            if np.any(r < 0):
                raise RuntimeError("HMatrix.truncate: block outside the kernel's capacity")
            new = blk.copy()
            new[lr, 4] = r
            old_dev = self._blk
            Un, Vn = self._factor_bufs(new)
            self._set_lowrank(new, Un, Vn)
            el = int(max((blk[lr, 1] * r).max(), (blk[lr, 3] * r).max()))
            zsplit = int(min(64, max(1, -(-el // 4096))))
            nat.call("mnpbem_hmat_trunc_form", code, nb, nat.ptr(old_dev), nat.ptr(self._blk), nat.ptr(bids),
                     nat.ptr(qoff), nat.ptr(Q), nat.ptr(RT), kst, zsplit, nat.ptr(Un), nat.ptr(Vn), st)
        self.htol = htol
        self._lu = None
        return self

    def _axpby(self, a, x, b=0.0, y=None, out=None):
        out = x if out is None else out
        ar, ai = complex(a).real, complex(a).imag
        br, bi = complex(b).real, complex(b).imag
        nat.call("mnpbem_hmat_axpby", _DT[self.dtype][1], x.numel(), ar, ai, nat.ptr(x), br, bi,
                 nat.ptr(y) if y is not None else None, nat.ptr(out), nat.stream_handle(self.device))
        self._pk = None
        return out

    def __rmul__(self, scalar):
        """scalar * H (hmatrix.py:553-559): dense blocks and left factors scaled."""
        if not isinstance(scalar, (int, float, complex, np.number)):
            return NotImplemented
        c = self._copy()
        if self._nval:
            c._axpby(scalar, c._val)
        if self._rsum:
            c._axpby(scalar, c._U)
        return c

    def __mul__(self, other):
        if isinstance(other, (int, float, complex, np.number)):
            return self.__rmul__(other)
        if isinstance(other, HMatrix):
            return self._mtimes_hmat(other)
        return self.mtimes_vec(other)

    def __neg__(self):
        return self.__rmul__(-1.0)

    def __add__(self, other):
        if not isinstance(other, HMatrix):
            raise TypeError("HMatrix: only an HMatrix can be added")
        return self._plus_hmat(other)

    def __sub__(self, other):
        return self.__add__(-other)

    def _plus_hmat(self, other):
        """Sum of two H-matrices on one block partition (hmatrix.py:592-670):
        dense blocks added, low-rank factors stacked [U1 U2], [V1 V2], then
        recompressed with truncate()."""
        self._check()
        other._check()
        a, b = self._blk_host, other._blk_host
        if a.shape != b.shape or not np.array_equal(a[:, :4], b[:, :4]):
            raise ValueError("HMatrix: the operands have different block partitions")
        if self.dtype != other.dtype:
            raise ValueError("HMatrix: the operands have different storage types")
        c = self._copy()
        if self._nval:
            c._axpby(1.0, self._val, 1.0, other._val, out=c._val)
        nd = self._nd
        if len(a) > nd:
            new = a.copy()
            new[nd:, 4] = np.maximum(a[nd:, 4], 0) + np.maximum(b[nd:, 4], 0)
            Un, Vn = c._factor_bufs(new)
            c._set_lowrank(new, Un, Vn)
            lr = np.nonzero(new[nd:, 4] > 0)[0] + nd
            if len(lr):
                bids = torch.as_tensor(lr.astype(np.int32)).to(self.device)
                nat.call("mnpbem_hmat_concat", _DT[self.dtype][1], len(lr), nat.ptr(self._blk), nat.ptr(other._blk),
                         nat.ptr(c._blk), nat.ptr(bids), nat.ptr(self._U), nat.ptr(self._V), nat.ptr(other._U),
                         nat.ptr(other._V), nat.ptr(Un), nat.ptr(Vn), nat.stream_handle(self.device))
        return c.truncate()

    def _diag_bids(self):
        blk = self._blk_host[:self._nd]
        ids = np.nonzero((blk[:, 0] == blk[:, 2]) & (blk[:, 1] == blk[:, 3]))[0]
        return torch.as_tensor(ids.astype(np.int32)).to(self.device), len(ids)

    def diag(self, xp=None):
        """Diagonal in face order (hmatrix.py:758-778), from the diagonal dense
        blocks: a complex128 device tensor, or numpy when xp is numpy."""
        self._check()
        n = self.tree.n
        out = torch.zeros((n,), dtype=torch.complex128, device=self.device)
        bids, nb = self._diag_bids()
        nat.call("mnpbem_hmat_diag", _DT[self.dtype][1], nb, nat.ptr(self._blk), nat.ptr(bids), nat.ptr(self._perm),
                 nat.ptr(self._val), nat.ptr(out), 0, nat.stream_handle(self.device))
        return out.cpu().numpy() if xp is np else out

    def pad(self):
        """Zero blocks for an H-matrix whose blocks were never filled
        (hmatrix.py:800-832): dense blocks zero, low-rank blocks zero rank-1."""
        if self._built:
            return self
        self._tables()
        blk = self._blk_host
        nd = self._nd
        tdt = _DT[self.dtype][0]
        self._val = torch.zeros((max(self._nval, 1),), dtype=tdt, device=self.device)
        blk[nd:, 4] = 1
        Un, Vn = self._factor_bufs(blk)
        Un.zero_()
        Vn.zero_()
        self._set_lowrank(blk, Un, Vn)
        self._built = True
        return self

    def eye_hmat(self):
        """Identity on this block partition (hmatrix.py:780-798): diagonal
        dense blocks are identities, everything else zero."""
        c = self._empty_like()
        c.pad()
        bids, nb = c._diag_bids()
        nat.call("mnpbem_hmat_diag", _DT[c.dtype][1], nb, nat.ptr(c._blk), nat.ptr(bids), nat.ptr(c._perm),
                 nat.ptr(c._val), None, 1, nat.stream_handle(c.device))
        c._pk = None
        return c

    def fillval(self, fun):
        """Refill the dense blocks from fun (hmatrix.py:834-852): an (n, n)
        matrix (device or host) or a callable fun(row, col) in face indices."""
        self._check()
        n = self.tree.n
        dev = self.device
        if callable(fun) and not isinstance(fun, (torch.Tensor, np.ndarray)):
            blk = self._blk_host[:self._nd]
            perm = self.tree.ind[:, 0]
            host = np.empty(max(self._nval, 1), dtype=np.complex128)
            for b in blk:
                rows = perm[b[0]:b[0] + b[1]]
                cols = perm[b[2]:b[2] + b[3]]
                cc, rr = np.meshgrid(cols, rows, indexing="ij")  # column-major block values
                host[b[5]:b[5] + b[1] * b[3]] = np.asarray(fun(rr.ravel(), cc.ravel()))
            self._val = torch.as_tensor(host).to(dev, _DT[self.dtype][0])
            self._lu = None
            self._pk = None
            return self
        src = torch.as_tensor(fun).to(dev, torch.complex128) if isinstance(fun, np.ndarray) else fun.to(torch.complex128)
        if tuple(src.shape) != (n, n):
            raise ValueError("HMatrix.fillval: expected an (n, n) matrix")
        si, sj = src.stride()
        dblk = torch.as_tensor(self._blk_host).to(dev)
        bids = torch.arange(self._nd, dtype=torch.int32, device=dev)
        nat.call("mnpbem_hmat_dense", _DT[self.dtype][1], self._nd, nat.ptr(dblk), nat.ptr(bids), nat.ptr(self._perm),
                 nat.ptr(src), si, sj, nat.ptr(self._val), nat.stream_handle(dev))
        self._lu = None
        self._pk = None
        return self

    # ------------------------------------------------------------ dense solves
    def _lu_dense(self, A):
        n = A.shape[0]
        dev = self.device
        piv = torch.empty((n,), dtype=torch.int32, device=dev)
        info = torch.empty((1,), dtype=torch.int32, device=dev)
        nat.call("mnpbem_getrf_batched", nat.DTYPE_C128, n, nat.ptr(A), n, n * n, nat.ptr(piv), n, nat.ptr(info), 1,
                 nat.stream_handle(dev))
        if int(info.cpu()[0]) != 0:
            raise np.linalg.LinAlgError("HMatrix: singular matrix (zero pivot in column %d)" % int(info.cpu()[0]))
        return A, piv

    def _getrs(self, lu, piv, B):
        n = lu.shape[0]
        nrhs = B.shape[1]
        nat.call("mnpbem_getrs_batched", nat.DTYPE_C128, n, nrhs, nat.ptr(lu), n, n * n, nat.ptr(piv), n,
                 nat.ptr(B), nrhs, n * nrhs, 1, nat.stream_handle(self.device))
        return B

    def lu(self, method="auto"):
        """LU factorisation for solve() (hmatrix.py:854-918).  The reference's
        'dense' and 'hierarchical' (2 x 2 Schur complement of the densified
        matrix) methods give the same solution up to rounding; both run here as
        the dense LU with partial pivoting on the GPU."""
        if method not in ("auto", "dense", "hierarchical"):
            raise ValueError("HMatrix.lu: method must be 'auto', 'dense' or 'hierarchical'")
        self._lu = self._lu_dense(self.full())
        self._lu_method = "dense"
        return self

    def solve(self, b):
        """x = A^-1 b for b (n,) or (n, k), numpy or torch (hmatrix.py:920-945);
        factorises on first use."""
        if getattr(self, "_lu", None) is None:
            self.lu()
        as_np = isinstance(b, np.ndarray)
        x = torch.as_tensor(b) if as_np else b
        one = x.dim() == 1
        X = x.reshape(x.shape[0], -1).to(self.device, torch.complex128).contiguous().clone()
        self._getrs(*self._lu, X)
        X = X[:, 0] if one else X
        return X.cpu().numpy() if as_np else X.to(x.device)

    def _empty_like(self):
        """Unfilled H-matrix with this tree, partition and settings."""
        c = HMatrix(None, htol=self.htol, kmax=self.kmax, dtype=self.dtype, device=self.device)
        c.tree = self.tree
        c.row1, c.col1, c.row2, c.col2 = self.row1, self.col1, self.row2, self.col2
        return c

    def _from_dense(self, M):
        """H-matrix on this partition compressed from the (n, n) device matrix M."""
        return self._empty_like().aca(M)

    def _dense_inv(self):
        lu, piv = self._lu_dense(self.full())
        n = self.tree.n
        E = torch.empty((n, n), dtype=torch.complex128, device=self.device)
        nat.call("mnpbem_identity_batched", nat.DTYPE_C128, n, nat.ptr(E), n, n * n, 1, nat.stream_handle(self.device))
        return self._getrs(lu, piv, E)

    def _gemm(self, A, B):
        """A @ B for (n, n) complex128 device matrices, on the matrix-core GEMM."""
        n = A.shape[0]
        C = torch.zeros((n, n), dtype=torch.complex128, device=self.device)
        negA = torch.empty_like(A)
        self._axpby_c128(-1.0, A, negA)
        nat.call("mnpbem_cgemm_sub_batched", nat.DTYPE_C128, n, n, n, nat.ptr(negA), n, n * n, nat.ptr(B), n, n * n,
                 nat.ptr(C), n, n * n, 1, nat.stream_handle(self.device))
        return C

    def _axpby_c128(self, a, x, out):
        nat.call("mnpbem_hmat_axpby", nat.DTYPE_C128, x.numel(), float(np.real(a)), float(np.imag(a)), nat.ptr(x),
                 0.0, 0.0, None, nat.ptr(out), nat.stream_handle(self.device))

    def _mtimes_hmat(self, other):
        """H @ H (hmatrix.py:947-965): product of the densified operands,
        compressed again by ACA on this partition."""
        return self._from_dense(self._gemm(self.full(), other.full()))

    def inv(self):
        """Inverse (hmatrix.py:982-996): dense LU inverse, compressed by ACA."""
        return self._from_dense(self._dense_inv())

    def lsolve(self, lu_hmat, key="N"):
        """X = A^-1 B with B = self, A = lu_hmat (hmatrix.py:998-1025), key 'N'."""
        if key != "N":
            raise NotImplementedError("HMatrix.lsolve: only key 'N' (full solve) is provided")
        lu, piv = lu_hmat._lu_dense(lu_hmat.full())
        return self._from_dense(self._getrs(lu, piv, self.full()))

    def rsolve(self, lu_hmat, key="N"):
        """X = B A^-1 with B = self, A = lu_hmat (hmatrix.py:1027-1056), key 'N'."""
        if key != "N":
            raise NotImplementedError("HMatrix.rsolve: only key 'N' (full solve) is provided")
        return self._from_dense(self._gemm(self.full(), lu_hmat._dense_inv()))


class ACACompGreenRet:
    """Green functions of a ComParticle compressed to H-matrices
    (aca_compgreen_ret.py:96-208).  ``eval(i, j, key, enei)`` returns an
    HMatrix of the CompGreenRet.eval matrix (or the int 0 for regions that no
    particle connects) whose entries the kernels generate in place, dense
    leaf blocks and ACA rows and columns alike, so no (n, n) matrix is
    assembled (mex/acagreen/acagreen.cpp:79-127); ``compress`` takes an
    assembled matrix instead.  The default admissibility is the
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
        self._partitions = {}

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

    def generate(self, side, key, kmat, enei):
        """HMatrix of side ``side``'s key ('G', 'F', 'H1', 'H2') at wavelength
        enei with entries generated in place; kmat: that wavelength's (nmat,
        2) row of CompGreenRet.wavenumbers."""
        # the four Green functions of one wavelength share the partition
        pkey = None if self._user_fadmiss is not None else float(enei)
        part = self._partitions.get(pkey)
        h = HMatrix(self.tree, htol=self.htol, kmax=self.kmax, fadmiss=self.fadmiss_for(enei), dtype=self.dtype,
                    device=self.g.device, partition=part)
        if part is None:
            if len(self._partitions) >= 8:
                self._partitions.clear()
            self._partitions[pkey] = h.partition
        return h.aca_green(self.g, side, key, kmat)

    def eval(self, i, j, key, enei):
        if key not in ("G", "F", "H1", "H2"):
            raise ValueError("ACACompGreenRet.eval: key must be G, F, H1 or H2")
        ck = (i, j, key, float(enei))
        if ck in self._cache:
            return self._cache[ck]
        if not np.any(self.g.con[i][j]):
            return 0
        if i != j:
            raise NotImplementedError("ACACompGreenRet.eval: cross-region blocks (nested particles)")
        h = self.generate(i, key, self.g.wavenumbers([enei])[0], enei)
        self._cache[ck] = h
        return h

    def eval_full(self, i, j, key, enei):
        h = self.eval(i, j, key, enei)
        return h if isinstance(h, int) else h.full()

    def compression(self, i=0, j=0, key="G", enei=600.0):
        return self.eval(i, j, key, enei).compression()

    def clear_cache(self):
        self._cache = {}
        self._partitions = {}
