"""GPU parity of the H-matrix recompression and algebra (hmatrix.py:543-1056):
truncate, scalar multiple, negation, sum and difference, diag, eye, pad,
fillval, LU solve, H @ H, inverse, lsolve / rsolve.

Oracles: truncate_block of the CPU restatement (oracle.hmatrix_oracle), run
on the very factors the GPU holds, pins the ranks and the recompressed blocks;
the SVD criterion itself is pinned on the CPU (tests/test_hmatrix_oracle.py).
The algebra is checked against numpy on the densified operands, with
tolerances written at each assertion: exact where the operation is exact
(negation, diag, eye), round-off where it is a dense solve, and the
compression tolerance where a result is ACA- or truncation-compressed.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle.hmatrix_oracle import truncate_block  # noqa: E402


def _cloud(n, seed, shift=0.0, radius=10.0):
    rng = np.random.RandomState(seed)
    ph = rng.uniform(0, 2 * np.pi, n)
    th = np.arccos(rng.uniform(-1, 1, n))
    p = radius * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    p[:, 0] += shift
    return p


def _helmholtz(pos, k=0.35):
    def fun(r, c):
        d = np.maximum(np.linalg.norm(pos[r] - pos[c], axis=1), 1e-10)
        return np.exp(1j * k * d) / (4 * np.pi * d)
    return fun


def _dense(fun, n):
    rr, cc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return fun(rr.ravel(), cc.ravel()).reshape(n, n)


def _hmat(pos, fun, htol, kmax, cleaf=16, dtype="complex128", ip=None):
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    t = ClusterTree(pos, cleaf=cleaf, ipart_arr=ip)
    return HMatrix.from_func(t, fun, htol=htol, kmax=kmax, dtype=dtype)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("htol", [1e-2, 1e-5, 1e-8])
def test_truncate_matches_oracle_blockwise(cuda, dtype, htol):
    """Each recompressed block equals truncate_block of the same factors:
    same rank, same product (to FP64 round-off for complex128 storage, to
    complex64 round-off otherwise)."""
    print("truncate", dtype, htol, flush=True)
    pos = _cloud(900, 1)
    h = _hmat(pos, _helmholtz(pos), 1e-11, 60, dtype=dtype)
    U0, V0 = h.lhs, h.rhs
    assert max(u.shape[1] for u in U0) > 8
    A0 = h.full(np)
    h.truncate(htol)
    U1, V1 = h.lhs, h.rhs
    tol = 1e-12 if dtype == "complex128" else 2e-6
    rank_mismatch = 0
    for u0, v0, u1, v1 in zip(U0, V0, U1, V1):
        if u0.shape[1] == 0:
            assert u1.shape[1] == 0
            continue
        uo, vo = truncate_block(u0, v0, htol)
        if u1.shape[1] != uo.shape[1]:
            # only a cumulative sum sitting on the threshold may round differently
            rank_mismatch += 1
            assert abs(u1.shape[1] - uo.shape[1]) == 1
            continue
        B = u0 @ v0.T
        assert np.linalg.norm(u1 @ v1.T - uo @ vo.T) <= tol * np.linalg.norm(B)
    assert rank_mismatch <= max(1, len(U0) // 100)
    assert h.ranks.sum() <= sum(u.shape[1] for u in U0)
    # the recompressed matrix stays within the truncation tolerance of the original
    assert _rel(h.full(np), A0) < 30 * htol + (1e-6 if dtype == "complex64" else 1e-12)


def test_truncate_rule_on_known_spectrum(cuda):
    """Two particles whose coupling block is B = P diag(s) W^T with known
    singular values s (P, W orthonormal): after ACA and truncate(htol) the
    block must be the best rank-r approximation of B, r being the reference's
    rule r = #{cumsum(s) < (1 - htol) sum(s)} (hmatrix.py:716-731; it drops
    the value that crosses the threshold, so a block of exact rank 2 keeps
    rank 1)."""
    print("truncate known spectrum", flush=True)
    rng = np.random.RandomState(12)
    n1, n2 = 300, 200
    pos = np.concatenate([_cloud(n1, 2), _cloud(n2, 9, shift=100.0)])
    ip = np.repeat([0, 1], [n1, n2])
    sv = 100.0 * np.array([1.0, 0.3, 0.1, 1e-4, 1e-7])
    P = np.linalg.qr(rng.randn(n1, 5) + 1j * rng.randn(n1, 5))[0]
    W = np.linalg.qr(rng.randn(n2, 5) + 1j * rng.randn(n2, 5))[0]
    B = (P * sv) @ W.T
    A = np.eye(n1 + n2, dtype=complex)
    A[:n1, n1:] = B
    A[n1:, :n1] = 0.5 * B.T
    cases = {1e-2: 2, 1e-5: 3, 1e-9: 4}
    for htol, r in cases.items():
        assert r == np.count_nonzero(np.cumsum(sv) < (1 - htol) * sv.sum())
        h = _hmat(pos, lambda i, j: A[i, j], 1e-14, 30, cleaf=20, ip=ip)
        h.truncate(htol)
        F = h.full(np)
        Br = (P[:, :r] * sv[:r]) @ W[:, :r].T
        assert np.linalg.norm(F[:n1, n1:] - Br) < 1e-11 * np.linalg.norm(B), htol
        assert np.linalg.norm(F[n1:, :n1] - 0.5 * Br.T) < 1e-11 * np.linalg.norm(B), htol
        assert h.ranks.max() == r


def test_truncate_high_rank_global_workspace(cuda):
    """Ranks above 64 take the global-workspace path of the SVD kernel."""
    print("truncate rank > 64", flush=True)
    pos = np.concatenate([_cloud(700, 3, radius=12.0), _cloud(700, 4, shift=80.0, radius=12.0)])
    ip = np.repeat([0, 1], 700)
    fun = _helmholtz(pos, k=1.2)
    h = _hmat(pos, fun, 1e-13, 110, cleaf=400, ip=ip)
    assert h.ranks.max() > 64, h.ranks.max()
    U0, V0 = h.lhs, h.rhs
    h.truncate(1e-9)
    for u0, v0, u1, v1 in zip(U0, V0, h.lhs, h.rhs):
        if u0.shape[1] <= 64:
            continue
        uo, vo = truncate_block(u0, v0, 1e-9)
        assert abs(u1.shape[1] - uo.shape[1]) <= 1
        assert np.linalg.norm(u1 @ v1.T - u0 @ v0.T) <= 1e-7 * np.linalg.norm(u0 @ v0.T)


def test_scalar_negation_sum_difference(cuda):
    """hmatrix.py:553-670 and the reference's tests/test_hmatrix.py
    (TestHMatrixMultiply.test_scalar_multiply, TestHMatrixOperations)."""
    print("hmatrix algebra", flush=True)
    pos = _cloud(400, 5)
    fun = _helmholtz(pos)
    h = _hmat(pos, fun, 1e-6, 40)
    A = h.full(np)
    v = np.random.RandomState(55).randn(400)
    # scaling rounds each product once: FP64 round-off
    assert _rel((3.0 * h) @ v, 3.0 * (h @ v)) < 1e-14
    assert _rel(((2 - 1j) * h).full(np), (2 - 1j) * A) < 1e-14
    assert np.array_equal((-h).full(np), -A)
    # sums are recompressed at htol = 1e-6: the dropped singular values of a
    # block sum to at most htol times their total, i.e. a relative Frobenius
    # error below htol sqrt(rank) <= 7e-6 per block
    s = h + h
    assert _rel(s.full(np), 2 * A) < 1e-5
    # the singular values of 2 B are twice those of B and the rule is scale
    # free: the sum has the ranks of one copy truncated on its own
    ht = h._copy().truncate()
    assert np.max(np.abs(s.ranks - ht.ranks)) <= 1
    d = h - h
    assert np.linalg.norm(d.full(np)) < 1e-12 * np.linalg.norm(A)
    g = _hmat(pos, _helmholtz(pos, k=0.1), 1e-6, 40)
    B = g.full(np)
    assert _rel((h - 0.5 * g).full(np), A - 0.5 * B) < 2e-5
    # the operands are left unchanged
    assert np.array_equal(h.full(np), A)


def test_sum_with_ranks_above_64(cuda):
    """h + h with ranks above 32 stacks factors beyond 64 columns."""
    print("hmatrix sum rank > 64", flush=True)
    pos = np.concatenate([_cloud(500, 6, radius=12.0), _cloud(500, 7, shift=80.0, radius=12.0)])
    ip = np.repeat([0, 1], 500)
    h = _hmat(pos, _helmholtz(pos, k=1.0), 1e-10, 60, cleaf=300, ip=ip)
    assert h.ranks.max() > 32
    A = h.full(np)
    s = h + h
    assert _rel(s.full(np), 2 * A) < 1e-8


def test_diag_eye_pad_fillval(cuda):
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    print("diag eye pad fillval", flush=True)
    pos = _cloud(300, 8)
    fun = _helmholtz(pos)
    h = _hmat(pos, fun, 1e-8, 40, cleaf=10)
    A = h.full(np)
    assert np.array_equal(h.diag(np), np.diag(A))
    e = h.eye_hmat()
    assert np.array_equal(e.full(np), np.eye(300))
    p = HMatrix(ClusterTree(pos, cleaf=10), htol=1e-6).pad()
    assert not np.any(p.full(np))
    assert all(not np.any(v) for v in p.val)
    # fillval replaces the dense blocks only
    M = _dense(fun, 300)
    q = p.fillval(lambda r, c: 2.0 * M[r, c])
    Q = q.full(np)
    dense_mask = np.abs(Q) > 0
    assert np.array_equal(Q[dense_mask], 2.0 * M[dense_mask])
    q2 = HMatrix(ClusterTree(pos, cleaf=10), htol=1e-6).pad().fillval(torch.as_tensor(2.0 * M).cuda())
    assert np.array_equal(q2.full(np), Q)


def test_lu_solve_inverse_products(cuda):
    """LU solve against numpy's dense solve (round-off), inverse and H @ H
    against their dense counterparts (compression tolerance); the reference's
    TestHMatrixSolve uses the same diagonally shifted kernel."""
    print("hmatrix lu solve inv mtimes", flush=True)
    n = 400
    pos = _cloud(n, 10)
    base = _helmholtz(pos)

    def fun(r, c):
        v = base(r, c)
        return v + 10.0 * (r == c)

    h = _hmat(pos, fun, 1e-10, 60)
    A = h.full(np)
    b = np.random.RandomState(44).randn(n, 3) + 1j * np.random.RandomState(45).randn(n, 3)
    x = h.lu().solve(b)
    assert _rel(x, np.linalg.solve(A, b)) < 1e-12
    x1 = h.solve(b[:, 0])
    assert x1.shape == (n,) and _rel(x1, np.linalg.solve(A, b[:, 0])) < 1e-12
    Ainv = np.linalg.inv(A)
    assert _rel(h.inv().full(np), Ainv) < 1e-8
    assert _rel((h @ h).full(np), A @ A) < 1e-8
    g = _hmat(pos, _helmholtz(pos, k=0.2), 1e-10, 60)
    B = g.full(np)
    assert _rel(g.lsolve(h).full(np), np.linalg.solve(A, B)) < 1e-8
    assert _rel(g.rsolve(h).full(np), B @ Ainv) < 1e-8
    with pytest.raises(NotImplementedError):
        g.lsolve(h, key="L")
