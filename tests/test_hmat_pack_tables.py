"""CPU check of the packed product tables of the H-matrix (host logic, no GPU).

`pack_tables` describes two copies of the block storage (column panels of the
V factors, row panels of the dense blocks and U factors) and the items the
product kernels run over.  Here the pack and the product are replayed in numpy
exactly as the kernels read the tables, on block tables of real cluster trees
with random block values and random ranks (zero ranks included), and the result
is compared with the product of the matrix assembled block by block from the
block storage: a wrong offset, stride or index in the tables fails it.
"""

import numpy as np
import pytest

from mnpbem_amd.hmatrix import ClusterTree, HMatrix, pack_tables


def _sphere(n, seed):
    rng = np.random.RandomState(seed)
    ph = rng.uniform(0, 2 * np.pi, n)
    th = np.arccos(rng.uniform(-1, 1, n))
    return 10.0 * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)


def _random_storage(h, seed):
    """Block table with random ranks (as HMatrix._fill lays it out) and
    random dense values and factors."""
    h._tables()
    blk = h._blk_host.copy()
    nd = h._nd
    rng = np.random.RandomState(seed)
    m, nn = blk[nd:, 1], blk[nd:, 3]
    ranks = rng.randint(0, 6, size=len(m))
    blk[nd:, 4] = ranks
    blk[nd:, 5] = np.r_[0, np.cumsum(m * ranks)[:-1]]
    blk[nd:, 6] = np.r_[0, np.cumsum(nn * ranks)[:-1]]
    blk[nd:, 7] = np.r_[0, np.cumsum(ranks)[:-1]]
    cplx = lambda k: rng.randn(k) + 1j * rng.randn(k)
    val = cplx(max(h._nval, 1))
    U = cplx(max(int((m * ranks).sum()), 1))
    V = cplx(max(int((nn * ranks).sum()), 1))
    return blk, val, U, V, int(ranks.sum())


def _dense_from_blocks(blk, nd, n, val, U, V):
    M = np.zeros((n, n), complex)
    for b in blk:
        r0, m, c0, nn, k, oa, ob = b[:7]
        if k < 0:
            M[r0:r0 + m, c0:c0 + nn] = val[oa:oa + m * nn].reshape(nn, m).T
        elif k > 0:
            Ub = U[oa:oa + m * k].reshape(k, m).T
            Vb = V[ob:ob + nn * k].reshape(k, nn).T
            M[r0:r0 + m, c0:c0 + nn] = Ub @ Vb.T
    return M


def _replay(blk, nd, n, leafs, lst, val, U, V, rsum, perm, X):
    cdesc, citems, rdesc, ritems, (nvs, ntix, nP, ng) = pack_tables(blk, nd, n, leafs, lst)
    VS = np.full(max(nvs, 1), np.nan, complex)
    tix = np.full(max(ntix, 1), -1, np.int64)
    for dst, K, nn, ob, rank, tdst, trow, _ in cdesc:  # k_hmat_pack_cols
        for r in range(rank):
            VS[dst + np.arange(nn) * K + r] = V[ob + r * nn:ob + (r + 1) * nn]
            tix[tdst + r] = trow + r
    P = np.full(max(nP, 1), np.nan, complex)
    gidx = np.full(max(ng, 1), -1, np.int64)
    for dst, mi, nc, src, ld, dense, gdst, gstart in rdesc:  # k_hmat_pack_rows
        S = val if dense else U
        for q in range(nc):
            P[dst + q * mi:dst + (q + 1) * mi] = S[src + q * ld:src + q * ld + mi]
            gidx[gdst + q] = gstart + q
    assert not np.isnan(VS[:nvs]).any() and not np.isnan(P[:nP]).any(), "pack leaves holes"
    assert (tix[:ntix] >= n).all() and (gidx[:ng] >= 0).all()
    R = np.zeros((n + rsum, X.shape[1]), complex)
    R[:n] = X[perm]  # k_hmat_xperm
    for vs, K, o0, nn, c0, og in citems:  # k_hmat_vt_pk
        o = np.arange(o0, min(o0 + 64, K))
        panel = VS[vs:vs + nn * K].reshape(nn, K)
        R[tix[og + o]] = panel[:, o].T @ R[c0:c0 + nn]
    Y = np.full(X.shape, np.nan, complex)
    for p, mi, K, g, r0, _ in ritems:  # k_hmat_rows_pk
        A = P[p:p + mi * K].reshape(K, mi).T
        Y[perm[r0:r0 + mi]] = A @ R[gidx[g:g + K]]
    return Y, (len(citems), len(ritems))


@pytest.mark.parametrize("npts,cleaf,two", [(700, 16, False), (700, 40, True), (3000, 100, False)])
def test_packed_tables_replay_equals_block_product(npts, cleaf, two):
    pos = _sphere(npts, 3)
    ip = None
    if two:
        p2 = _sphere(500, 4)
        p2[:, 0] += 24.0
        pos = np.concatenate([pos, p2])
        ip = np.repeat([0, 1], [npts, 500])
    t = ClusterTree(pos, cleaf=cleaf, ipart_arr=ip)
    h = HMatrix(t, device="cpu")
    blk, val, U, V, rsum = _random_storage(h, seed=cleaf)
    n = t.n
    perm = t.ind[:, 0]
    X = np.random.RandomState(5).randn(n, 3) + 1j * np.random.RandomState(6).randn(n, 3)
    Y, (nci, nri) = _replay(blk, h._nd, n, h._leafs_host, h._lst_host, val, U, V, rsum, perm, X)
    M = _dense_from_blocks(blk, h._nd, n, val, U, V)
    Yref = np.empty_like(X)
    Yref[perm] = M @ X[perm]
    assert not np.isnan(Y).any(), "a row of Y is never written"
    assert np.linalg.norm(Y - Yref) <= 1e-12 * np.linalg.norm(Yref)
    assert (blk[h._nd:, 4] > 0).any() and nci > 0 and nri >= len(h._leafs_host)
    if cleaf > 64:  # leaves above 64 rows split into 64-row groups
        assert nri > len(h._leafs_host)


def test_packed_tables_all_dense_and_rank_zero():
    """No low-rank outputs at all (every admissible block of rank 0): no
    column-panel items, and the row panels hold only the dense blocks."""
    pos = _sphere(300, 8)
    t = ClusterTree(pos, cleaf=32)
    h = HMatrix(t, device="cpu")
    h._tables()
    blk = h._blk_host.copy()
    nd = h._nd
    blk[nd:, 4] = 0
    blk[nd:, 5:8] = 0
    cdesc, citems, rdesc, ritems, (nvs, ntix, nP, ng) = pack_tables(blk, nd, t.n, h._leafs_host, h._lst_host)
    assert len(cdesc) == 0 and len(citems) == 0 and nvs == 0 and ntix == 0
    assert nP == h._nval and len(rdesc) == nd
