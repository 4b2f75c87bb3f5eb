"""GPU parity of the dense complex kernels (MFMA GEMM, batched LU, LU
solves) against numpy on the same seeded inputs, through the C ABI."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _lib():
    from mnpbem_amd import _native as nat
    return nat


def _rand(rng, shape, dtype):
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return a.astype(dtype)


@pytest.mark.parametrize("dtype,tol,split", [(np.complex128, 1e-13, False), (np.complex64, 2e-5, False),
                                             (np.complex64, 1e-4, True)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 37, 53), (300, 200, 129), (1, 1, 1), (17, 130, 5),
                                   (257, 129, 1000)])
def test_cgemm_sub(cuda, dtype, tol, split, M, N, K):
    nat = _lib()
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    B = 3
    A = _rand(rng, (B, M, K), dtype)
    Bm = _rand(rng, (B, K, N), dtype)
    C = _rand(rng, (B, M, N), dtype)
    ref = C.astype(np.complex128) - A.astype(np.complex128) @ Bm.astype(np.complex128)
    tA, tB, tC = (torch.from_numpy(x).to(cuda) for x in (A, Bm, C))
    dt = nat.DTYPE_C128 if dtype == np.complex128 else (nat.DTYPE_C64X if split else nat.DTYPE_C64)
    print("cgemm", dtype.__name__, split, M, N, K, flush=True)
    nat.call("mnpbem_cgemm_sub_batched", dt, M, N, K, nat.ptr(tA), K, M * K, nat.ptr(tB), N, K * N, nat.ptr(tC), N,
             M * N, B, nat.stream_handle(cuda))
    out = tC.cpu().numpy()
    err = np.abs(out - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("M,N,K", [(2884, 8, 8), (1000, 16, 16), (300, 1, 64), (5, 16, 3), (700, 13, 40),
                                   (2884, 32, 32), (129, 31, 64)])
def test_thin_products_fp32(cuda, M, N, K):
    """Complex64 products with N <= 32 and K <= 64 (the bottom of the LU
    recursion) on the FP32 vector kernel (mnpbem_gemm_thin_max, built in for
    N <= 16): against numpy to FP32 accuracy, with strided operands and the
    padding of C untouched."""
    nat = _lib()
    rng = np.random.default_rng(M + 3 * N + 7 * K)
    B = 3
    lda, ldb, ldc = K + 2, N + 1, N + 5
    A = _rand(rng, (B, M, lda), np.complex64)
    Bm = _rand(rng, (B, K, ldb), np.complex64)
    C = _rand(rng, (B, M, ldc), np.complex64)
    ref = C.astype(np.complex128).copy()
    ref[:, :, :N] -= A[:, :, :K].astype(np.complex128) @ Bm[:, :, :N].astype(np.complex128)
    tA, tB, tC = (torch.from_numpy(x).to(cuda) for x in (A, Bm, C))
    print("thin product", M, N, K, flush=True)
    lib = nat.load()
    prev = lib.mnpbem_gemm_thin_max(32)  # every N <= 32 on the FP32 kernel (built-in: N <= 16)
    try:
        nat.call("mnpbem_cgemm_sub_batched", nat.DTYPE_C64X, M, N, K, nat.ptr(tA), lda, M * lda, nat.ptr(tB), ldb,
                 K * ldb, nat.ptr(tC), ldc, M * ldc, B, nat.stream_handle(cuda))
    finally:
        lib.mnpbem_gemm_thin_max(prev)
    out = tC.cpu().numpy()
    scale = np.abs(ref[:, :, :N]).max() + np.sqrt(K) * 2.0
    assert np.abs(out[:, :, :N] - ref[:, :, :N]).max() / scale < 3e-6
    assert np.array_equal(out[:, :, N:], C[:, :, N:])


@pytest.mark.parametrize("M,N,K", [(130, 140, 200), (300, 129, 96), (64, 64, 33)])
def test_bf16x3_strided_operands(cuda, M, N, K):
    """Operands that are views into larger matrices (leading dimensions above
    the row length, odd ones included, batch strides above the matrix size):
    the bf16x3 GEMM against numpy, with the padding of C left untouched."""
    nat = _lib()
    rng = np.random.default_rng(M + N + K)
    B = 2
    lda, ldb, ldc = K + 3, N + 5, N + 7
    A = _rand(rng, (B, M + 1, lda), np.complex64)
    Bm = _rand(rng, (B, K + 2, ldb), np.complex64)
    C = _rand(rng, (B, M + 2, ldc), np.complex64)
    ref = C.astype(np.complex128).copy()
    ref[:, :M, :N] -= A[:, :M, :K].astype(np.complex128) @ Bm[:, :K, :N].astype(np.complex128)
    tA, tB, tC = (torch.from_numpy(x).to(cuda) for x in (A, Bm, C))
    print("bf16x3 strided", M, N, K, flush=True)
    for v in (0, 1, 2):
        tC.copy_(torch.from_numpy(C))
        nat.call("mnpbem_cgemm_sub_variant", v, M, N, K, nat.ptr(tA), lda, (M + 1) * lda, nat.ptr(tB), ldb,
                 (K + 2) * ldb, nat.ptr(tC), ldc, (M + 2) * ldc, B, nat.stream_handle(cuda))
        out = tC.cpu().numpy()
        err = np.abs(out[:, :M, :N] - ref[:, :M, :N]).max() / np.abs(ref[:, :M, :N]).max()
        assert err < 1e-4, (v, err)
        assert np.array_equal(out[:, M:, :], C[:, M:, :]) and np.array_equal(out[:, :, N:], C[:, :, N:])


@pytest.mark.parametrize("M,N,K", [(64, 300, 64), (130, 129, 200), (17, 64, 33), (200, 70, 96), (1, 5, 1),
                                   (257, 300, 161), (129, 131, 224), (65, 66, 97), (40, 50, 31), (128, 128, 128)])
def test_bf16x3_tile_variants_identical(cuda, M, N, K):
    """The three tile variants of the bf16x3 GEMM (and the table-routed call)
    give bit-identical results: each output element is accumulated over K in
    the same order.  All agree with numpy to the split-product accuracy."""
    nat = _lib()
    from mnpbem_amd import autotune as at
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    B = 2
    A = _rand(rng, (B, M, K), np.complex64)
    Bm = _rand(rng, (B, K, N), np.complex64)
    C = _rand(rng, (B, M, N), np.complex64)
    ref = C.astype(np.complex128) - A.astype(np.complex128) @ Bm.astype(np.complex128)
    tA, tB = torch.from_numpy(A).to(cuda), torch.from_numpy(Bm).to(cuda)
    st = nat.stream_handle(cuda)
    print("bf16x3 variants", M, N, K, flush=True)
    outs = []
    for v in (0, 1, 2):
        tC = torch.from_numpy(C).to(cuda)
        nat.call("mnpbem_cgemm_sub_variant", v, M, N, K, nat.ptr(tA), K, M * K, nat.ptr(tB), N, K * N, nat.ptr(tC),
                 N, M * N, B, st)
        outs.append(tC.cpu().numpy())
    saved = at.get_table()
    # products with N <= 16 and K <= 64 leave the table for the FP32 kernel
    thin = N <= 16 and K <= 64
    try:
        for v in ((0, 1, 2) if not thin else ()):
            at.set_table(np.full(at.NCELL, v, dtype=np.uint8))
            tC = torch.from_numpy(C).to(cuda)
            nat.call("mnpbem_cgemm_sub_batched", nat.DTYPE_C64X, M, N, K, nat.ptr(tA), K, M * K, nat.ptr(tB), N,
                     K * N, nat.ptr(tC), N, M * N, B, st)
            outs.append(tC.cpu().numpy())
    finally:
        at.set_table(saved)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert np.abs(outs[0] - ref).max() / np.abs(ref).max() < 1e-4


def test_autotune_records_and_installs(cuda):
    """tune() times every variant on the recorded shapes and installs the
    faster one per class; the recorded shape is the launched one."""
    nat = _lib()
    from mnpbem_amd import autotune as at
    M, N, K, B = 64, 320, 64, 4
    tA = torch.randn((B, M, K), dtype=torch.complex64, device=cuda)
    tB = torch.randn((B, K, N), dtype=torch.complex64, device=cuda)
    tC = torch.zeros((B, M, N), dtype=torch.complex64, device=cuda)
    st = nat.stream_handle(cuda)

    def run():
        nat.call("mnpbem_cgemm_sub_batched", nat.DTYPE_C64X, M, N, K, nat.ptr(tA), K, M * K, nat.ptr(tB), N, K * N,
                 nat.ptr(tC), N, M * N, B, st)

    saved = at.get_table()
    try:
        table, report = at.tune(run, reps=1, verbose=False)
        assert [r[1] for r in report] == [(M, N, K, B)]
        cell, _, t, choice = report[0]
        assert cell == at.cell_of(M, N, K) and choice == int(np.argmin(t)) and all(x > 0 for x in t)
        assert at.get_table()[cell] == choice
    finally:
        at.set_table(saved)


@pytest.mark.parametrize("dtype,tol,split", [(np.complex128, 1e-12, False), (np.complex64, 1e-4, False),
                                             (np.complex64, 1e-3, True)])
@pytest.mark.parametrize("n", [1, 7, 64, 100, 300, 1031])
def test_getrf_getrs(cuda, dtype, tol, split, n):
    nat = _lib()
    rng = np.random.default_rng(n)
    B = 2
    A = (_rand(rng, (B, n, n), dtype) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(dtype)
    X = _rand(rng, (B, n, n + 3), dtype)
    ref = np.linalg.solve(A.astype(np.complex128), X.astype(np.complex128))
    tA, tX = torch.from_numpy(A.copy()).to(cuda), torch.from_numpy(X.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else (nat.DTYPE_C64X if split else nat.DTYPE_C64)
    st = nat.stream_handle(cuda)
    print("getrf", dtype.__name__, split, n, flush=True)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    nat.call("mnpbem_getrs_batched", dt, n, n + 3, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(tX), n + 3,
             n * (n + 3), B, st)
    assert np.all(info.cpu().numpy() == 0)
    err = np.abs(tX.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("dtype,tol,split", [(np.complex128, 1e-12, False), (np.complex64, 1e-4, False),
                                             (np.complex64, 1e-3, True)])
@pytest.mark.parametrize("n,nrhs", [(64, 5), (61, 64), (200, 130), (200, 1100), (45, 300), (300, 1), (30, 40),
                                    (17, 3)])
def test_trsm_row_blocked_leaves(cuda, dtype, tol, split, n, nrhs):
    """Triangular leaves with few right-hand sides run the row-blocked kernel
    (inside getrf's recursion, and in getrs up to 32768 columns over the
    batch): the factors and solutions match numpy's solve, and the solution
    matches the one-thread-per-column leaves (knob at 0) to rounding."""
    nat = _lib()
    lib = nat.load()
    rng = np.random.default_rng(n + nrhs)
    B = 2
    A = (_rand(rng, (B, n, n), dtype) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(dtype)
    X = _rand(rng, (B, n, nrhs), dtype)
    ref = np.linalg.solve(A.astype(np.complex128), X.astype(np.complex128))
    dt = nat.DTYPE_C128 if dtype == np.complex128 else (nat.DTYPE_C64X if split else nat.DTYPE_C64)
    st = nat.stream_handle(cuda)
    outs = []
    prev = lib.mnpbem_trsm_rows_max(-1)
    try:
        for knob in (-1, 0):
            lib.mnpbem_trsm_rows_max(knob)
            tA, tX = torch.from_numpy(A.copy()).to(cuda), torch.from_numpy(X.copy()).to(cuda)
            piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
            info = torch.empty((B,), dtype=torch.int32, device=cuda)
            print("trsm rows leaf", dtype.__name__, split, n, nrhs, knob, flush=True)
            nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
            nat.call("mnpbem_getrs_batched", dt, n, nrhs, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(tX), nrhs,
                     n * nrhs, B, st)
            assert np.all(info.cpu().numpy() == 0)
            outs.append(tX.cpu().numpy())
    finally:
        lib.mnpbem_trsm_rows_max(prev)
    for o in outs:
        assert np.abs(o - ref).max() / np.abs(ref).max() < tol
    assert np.abs(outs[0] - outs[1]).max() / np.abs(ref).max() < tol


def test_getrf_pivoting_matches_lapack(cuda):
    """Random (non-dominant) matrices force row interchanges; pivots must
    equal LAPACK zgetrf's (largest |re|+|im|, first on ties)."""
    import scipy.linalg as sla
    nat = _lib()
    rng = np.random.default_rng(11)
    n, B = 200, 2
    A = _rand(rng, (B, n, n), np.complex128)
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C128, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B,
             nat.stream_handle(cuda))
    out = tA.cpu().numpy()
    for b in range(B):
        lu, p = sla.lu_factor(A[b])
        assert np.array_equal(p, piv[b].cpu().numpy())
        assert np.abs(out[b] - lu).max() < 1e-11 * np.abs(lu).max()


def test_singular_matrix_reports_info(cuda):
    nat = _lib()
    n = 40
    A = np.eye(n, dtype=np.complex128)
    A[:, 17] = 0.0  # column 17 zero -> zero pivot at step 17 (1-based 18)
    tA = torch.from_numpy(A[None].copy()).to(cuda)
    piv = torch.empty((1, n), dtype=torch.int32, device=cuda)
    info = torch.empty((1,), dtype=torch.int32, device=cuda)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C128, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), 1,
             nat.stream_handle(cuda))
    assert int(info.cpu()[0]) == 18


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 1e-4)])
@pytest.mark.parametrize("n,C", [(5, 3), (64, 8), (130, 2), (700, 16), (701, 19)])
@pytest.mark.parametrize("blockinv", [False, True])
def test_lu_solve_transposed(cuda, dtype, tol, n, C, blockinv):
    """Solve M x = v from the LU of the buffer M^T (BEMRet's storage), with and
    without the precomputed inverse diagonal blocks."""
    nat = _lib()
    rng = np.random.default_rng(n + C)
    B = 2
    M = (_rand(rng, (B, n, n), dtype) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(dtype)
    V = _rand(rng, (B, n, C), dtype)
    ref = np.linalg.solve(M.astype(np.complex128), V.astype(np.complex128))
    tA = torch.from_numpy(np.ascontiguousarray(np.transpose(M, (0, 2, 1)))).to(cuda)
    tV = torch.from_numpy(V.copy()).to(cuda)
    tZ = torch.empty_like(tV)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else nat.DTYPE_C64
    st = nat.stream_handle(cuda)
    print("lu_solve_t", dtype.__name__, n, C, blockinv, flush=True)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    D, sD = None, 0
    if blockinv:
        sD = ((n + 63) // 64) * 2 * 64 * 64
        D = torch.empty((B, sD), dtype=tA.dtype, device=cuda)
        nat.call("mnpbem_lu_blockinv_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(D), sD, B, st)
    nat.call("mnpbem_lu_solve_t_batched", dt, n, C, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(tV), C, n * C,
             nat.ptr(tZ), C, n * C, nat.ptr(D), sD, B, st)
    err = np.abs(tV.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("dtype,tol,mf", [(np.complex128, 1e-11, 3), (np.complex64, 1e-4, 3),
                                          (np.complex64, 1e-4, 0)])
@pytest.mark.parametrize("n,C,B", [(700, 8, 200), (1000, 2, 130), (513, 5, 260), (1030, 6, 128)])
def test_lu_solve_transposed_many_systems(cuda, dtype, tol, mf, n, C, B):
    """Enough systems for the 256-row tiles of the windowed solve, where
    workgroup 0 of each launch solves the next window once and the next launch
    reads that solution (look-ahead), against numpy's solve; complex64 tile
    updates on the matrix cores (knob mnpbem_trsv_mfma 3, n even) and on the
    vector ALUs (0)."""
    nat = _lib()
    lib = nat.load()
    prev = lib.mnpbem_trsv_mfma(mf)
    print("lu_solve_t many: trsv_mfma", mf, flush=True)
    try:
        _solve_t_many(cuda, nat, dtype, tol, n, C, B)
    finally:
        lib.mnpbem_trsv_mfma(prev)


def _solve_t_many(cuda, nat, dtype, tol, n, C, B):
    rng = np.random.default_rng(n + C + B)
    M = (_rand(rng, (B, n, n), dtype) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(dtype)
    V = _rand(rng, (B, n, C), dtype)
    ref = np.linalg.solve(M.astype(np.complex128), V.astype(np.complex128))
    assert -(-n // 256) * B >= 512
    tA = torch.from_numpy(np.ascontiguousarray(np.transpose(M, (0, 2, 1)))).to(cuda)
    tV = torch.from_numpy(V.copy()).to(cuda)
    tZ = torch.empty_like(tV)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else nat.DTYPE_C64
    st = nat.stream_handle(cuda)
    print("lu_solve_t many", dtype.__name__, n, C, B, flush=True)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    sD = ((n + 63) // 64) * 2 * 64 * 64
    D = torch.empty((B, sD), dtype=tA.dtype, device=cuda)
    nat.call("mnpbem_lu_blockinv_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(D), sD, B, st)
    nat.call("mnpbem_lu_solve_t_batched", dt, n, C, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(tV), C, n * C,
             nat.ptr(tZ), C, n * C, nat.ptr(D), sD, B, st)
    err = np.abs(tV.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < tol, err


@pytest.mark.parametrize("dtype,n,tol", [(np.complex64, 14000, 2e-4), (np.complex128, 13000, 1e-11),
                                         (np.complex64, 6500, 1e-4)])
def test_getrf_tall_panels(cuda, dtype, n, tol):
    """Matrices taller than the 8-column register panel allows (one- and
    two-column panels for the first columns): solve A x = b and check the
    residual in FP64 on the device (diagonally weighted random A)."""
    nat = _lib()
    print("getrf tall", dtype.__name__, n, flush=True)
    g = torch.Generator(device=cuda).manual_seed(n)
    cdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    A = (torch.randn((n, n), dtype=cdt, device=cuda, generator=g) / np.sqrt(n)
         + 3.0 * torch.eye(n, dtype=cdt, device=cuda))
    x0 = torch.randn((n, 2), dtype=cdt, device=cuda, generator=g)
    b = (A.to(torch.complex128) @ x0.to(torch.complex128)).to(cdt)
    LU = A.clone()
    piv = torch.empty((n,), dtype=torch.int32, device=cuda)
    info = torch.empty((1,), dtype=torch.int32, device=cuda)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else nat.DTYPE_C64
    st = nat.stream_handle(cuda)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(LU), n, n * n, nat.ptr(piv), n, nat.ptr(info), 1, st)
    x = b.clone()
    nat.call("mnpbem_getrs_batched", dt, n, 2, nat.ptr(LU), n, n * n, nat.ptr(piv), n, nat.ptr(x), 2, 2 * n, 1, st)
    assert int(info.cpu()[0]) == 0
    p = piv.cpu().numpy()
    assert np.all(p >= np.arange(n)) and np.all(p < n)
    r = b.to(torch.complex128) - A.to(torch.complex128) @ x.to(torch.complex128)
    rel = (torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b.to(torch.complex128))).item()
    assert rel < tol, rel
    del A, LU, x, b




@pytest.mark.parametrize("dtype,n", [(np.complex128, 3500), (np.complex64, 4100)])
def test_getrf_tall_panels_match_lapack(cuda, dtype, n):
    """With the threshold lowered to 3072 rows (built in: 12288 complex64,
    6144 complex128) the panels are factored by the multi-workgroup column
    kernels (k_tall_pivot / k_tall_update): same pivots as LAPACK and factors
    within complex128 rounding; complex64 a valid pivoted LU; a zero column
    is reported in info."""
    import scipy.linalg as sla
    nat = _lib()
    # exercise the column kernels above 3072 rows (built in: 12288 / 6144);
    # the knob is process-wide, so it is restored whatever the outcome
    prev = nat.load().mnpbem_getrf_tall_rows(3072)
    try:
        _tall_panels_case(nat, sla, cuda, dtype, n)
    finally:
        nat.load().mnpbem_getrf_tall_rows(prev)


def _tall_panels_case(nat, sla, cuda, dtype, n):
    rng = np.random.default_rng(5)
    A = _rand(rng, (1, n, n), dtype)
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((1, n), dtype=torch.int32, device=cuda)
    info = torch.empty((1,), dtype=torch.int32, device=cuda)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else nat.DTYPE_C64
    print("getrf tall", dtype.__name__, n, flush=True)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), 1,
             nat.stream_handle(cuda))
    out = tA.cpu().numpy()[0]
    lu, p = sla.lu_factor(A[0].astype(np.complex128))
    got = piv[0].cpu().numpy()
    if dtype == np.complex128:
        assert np.array_equal(p, got)
        assert np.abs(out - lu).max() < 1e-10 * np.abs(lu).max()
    else:
        # complex64 rounding flips near-ties, after which the pivot sequence
        # departs from LAPACK's complex128 one: check the factorisation itself,
        # P A = L U to complex64 accuracy, and the partial-pivoting bound
        # |l| <= sqrt(2) (pivots chosen by the largest |re|+|im|)
        L = np.tril(out.astype(np.complex128), -1) + np.eye(n)
        U = np.triu(out.astype(np.complex128))
        PA = A[0].astype(np.complex128).copy()
        for i, q in enumerate(got):
            PA[[i, q]] = PA[[q, i]]
        assert np.abs(PA - L @ U).max() < 1e-4 * np.abs(A[0]).max()
        assert np.abs(L).max() <= np.sqrt(2.0) * (1 + 1e-5)
    assert int(info.cpu()[0]) == 0
    # singular: column 300 (inside the tall-panel range, c0 < n - 3072) zero -> info 301
    Z = np.eye(n, dtype=dtype)[None].copy()
    Z[0, :, 300] = 0.0
    tZ = torch.from_numpy(Z).to(cuda)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tZ), n, n * n, nat.ptr(piv), n, nat.ptr(info), 1,
             nat.stream_handle(cuda))
    assert int(info.cpu()[0]) == 301


def test_getri_above_16384_rows(cuda):
    """getri on one complex64 matrix of 17000 rows (the permutation of the
    pivots no longer fits 64 KB of LDS; the 25 000-face dimer needs such
    sizes): A^-1 (A x) against x for random vectors, with the products taken
    by torch on the GPU.  The rows of a diagonally dominant matrix are
    permuted at random, so the pivots undo that permutation and the column
    scatter moves entries across the whole 0..16999 range."""
    nat = _lib()
    n = 17000
    g = torch.Generator(device=cuda).manual_seed(5)
    A = torch.randn((n, n), dtype=torch.complex64, device=cuda, generator=g)
    A.diagonal().add_(4.0 * np.sqrt(n))
    A = A[torch.randperm(n, device=cuda, generator=g)].contiguous()
    tA = A.clone()
    piv = torch.empty((1, n), dtype=torch.int32, device=cuda)
    info = torch.empty((1,), dtype=torch.int32, device=cuda)
    W = torch.empty_like(tA)
    Q = torch.empty((1, n), dtype=torch.int32, device=cuda)
    st = nat.stream_handle(cuda)
    print("getri", n, flush=True)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), 1, st)
    nat.call("mnpbem_getri_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(W), n,
             n * n, nat.ptr(Q), n, 1, st)
    del W
    assert int(info.cpu()[0]) == 0
    pv = piv[0].cpu().numpy()
    assert (pv != np.arange(n)).sum() > n // 2 and pv.max() > 16384
    x = torch.randn((n, 3), dtype=torch.complex64, device=cuda, generator=g)
    y = tA @ (A @ x)
    err = ((y - x).abs().max() / x.abs().max()).item()
    assert err < 1e-3, err


@pytest.mark.parametrize("dtype,n", [(np.complex128, 300), (np.complex128, 517), (np.complex64, 300)])
def test_getri_matches_inverse(cuda, dtype, n):
    """mnpbem_getri_batched (triangular inverse of L, dense upper solve, column
    scatter by the pivots) against numpy's inverse; complex64 through the
    bf16x3 GEMMs to the split-product accuracy."""
    nat = _lib()
    rng = np.random.default_rng(n)
    B = 2
    A = _rand(rng, (B, n, n), dtype)
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    W = torch.empty_like(tA)
    Q = torch.empty((B, n), dtype=torch.int32, device=cuda)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else nat.DTYPE_C64X
    st = nat.stream_handle(cuda)
    print("getri", dtype.__name__, n, flush=True)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    nat.call("mnpbem_getri_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(W), n, n * n,
             nat.ptr(Q), n, B, st)
    out = tA.cpu().numpy()
    # the inverse the solver formed before: a solve against the identity
    tB = torch.from_numpy(A.copy()).to(cuda)
    nat.call("mnpbem_getrf_batched", dt, n, nat.ptr(tB), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    E = torch.zeros_like(tB)
    E[:] = torch.eye(n, dtype=tB.dtype, device=cuda)
    nat.call("mnpbem_getrs_batched", dt, n, n, nat.ptr(tB), n, n * n, nat.ptr(piv), n, nat.ptr(E), n, n * n, B, st)
    old = E.cpu().numpy()
    for b in range(B):
        ref = np.linalg.inv(A[b].astype(np.complex128))
        err = np.abs(out[b] - ref).max() / np.abs(ref).max()
        err_old = np.abs(old[b] - ref).max() / np.abs(ref).max()
        if dtype == np.complex128:
            assert err < 1e-11, (b, err)
        else:
            # complex64 / bf16x3: the condition number (~n for random
            # matrices) times the product accuracy; no worse than before
            assert err < 3 * err_old + 1e-6, (b, err, err_old)


@pytest.mark.parametrize("bs", [64, 128])
@pytest.mark.parametrize("n,nrhs", [(1, 3), (7, 70), (64, 64), (65, 129), (130, 40), (300, 303), (1031, 200),
                                    (2884, 64)])
def test_getrs_blockinv_leaves(cuda, n, nrhs, bs):
    """mnpbem_getrs_blockinv_batched: the 64-row leaves as products with the
    inverse diagonal blocks (bf16x3) against numpy's solve, and no less
    accurate than the substitution leaves of mnpbem_getrs_batched; the
    partial last leaf (n mod 64 rows) included.  Other dtypes fall through
    to the plain entry point (bit-identical)."""
    nat = _lib()
    rng = np.random.default_rng(n + nrhs)
    B = 2
    A = (_rand(rng, (B, n, n), np.complex64) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(np.complex64)
    X = _rand(rng, (B, n, nrhs), np.complex64)
    ref = np.linalg.solve(A.astype(np.complex128), X.astype(np.complex128))
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    st = nat.stream_handle(cuda)
    print("getrs_blockinv", n, nrhs, flush=True)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    D, sD = _block_inverses(nat, tA, n, B, bs, st)
    out = {}
    for name in ("plain", "blockinv", "fp32"):
        tX = torch.from_numpy(X.copy()).to(cuda)
        if name == "plain":
            nat.call("mnpbem_getrs_batched", nat.DTYPE_C64X, n, nrhs, nat.ptr(tA), n, n * n, nat.ptr(piv), n,
                     nat.ptr(tX), nrhs, n * nrhs, B, st)
        else:
            dt = nat.DTYPE_C64X if name == "blockinv" else nat.DTYPE_C64
            nat.call("mnpbem_getrs_blockinv_batched", dt, n, nrhs, nat.ptr(tA), n, n * n, nat.ptr(piv), n,
                     nat.ptr(D), sD, bs, nat.ptr(tX), nrhs, n * nrhs, B, st)
        out[name] = tX.cpu().numpy()
    assert np.all(info.cpu().numpy() == 0)
    scale = np.abs(ref).max()
    err = np.abs(out["blockinv"] - ref).max() / scale
    err_plain = np.abs(out["plain"] - ref).max() / scale
    # a leaf product in bf16x3 is accurate to ~2^-16 of |D| |X| (the
    # substitution it replaces runs in FP32: exact for n = 1)
    assert err < 3 * err_plain + 1e-4, (err, err_plain)
    # complex64 without the split products ignores the inverses
    tX = torch.from_numpy(X.copy()).to(cuda)
    nat.call("mnpbem_getrs_batched", nat.DTYPE_C64, n, nrhs, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(tX),
             nrhs, n * nrhs, B, st)
    assert np.array_equal(tX.cpu().numpy(), out["fp32"])


@pytest.mark.parametrize("n,nrhs", [(300, 512), (1031, 700), (2884, 600)])
def test_getrs_getri_left_looking(cuda, n, nrhs):
    """The many-column solves with 128-row inverse blocks block row by block
    row (mnpbem_trsm_left_rhs, from 512 right-hand sides) against numpy's
    solve and inverse, as accurate as the recursive solves (knob 0) on the
    same factors; the partial last block row included.  The leaf folded into
    the update product (mnpbem_trsm_leaf_fused) gives the bits of the
    two-launch form."""
    nat = _lib()
    rng = np.random.default_rng(5 * n + nrhs)
    B = 2
    A = (_rand(rng, (B, n, n), np.complex64) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(np.complex64)
    X = _rand(rng, (B, n, nrhs), np.complex64)
    ref = np.linalg.solve(A.astype(np.complex128), X.astype(np.complex128))
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    st = nat.stream_handle(cuda)
    print("getrs/getri left-looking", n, nrhs, flush=True)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    D, sD = _block_inverses(nat, tA, n, B, 128, st)
    W = torch.empty_like(tA)
    Q = torch.empty((B, n), dtype=torch.int32, device=cuda)
    lib = nat.load()
    out, inv = {}, {}
    prev = lib.mnpbem_trsm_left_rhs(-1)
    prev_f = lib.mnpbem_trsm_leaf_fused(1)
    try:
        # 0: recursive; 512: block rows with the leaf folded into the update
        # product; 513: block rows, update and leaf as two launches
        for knob in (0, 512, 513):
            lib.mnpbem_trsm_left_rhs(min(knob, 512))
            lib.mnpbem_trsm_leaf_fused(0 if knob == 513 else 1)
            tX = torch.from_numpy(X.copy()).to(cuda)
            nat.call("mnpbem_getrs_blockinv_batched", nat.DTYPE_C64X, n, nrhs, nat.ptr(tA), n, n * n, nat.ptr(piv),
                     n, nat.ptr(D), sD, 128, nat.ptr(tX), nrhs, n * nrhs, B, st)
            out[knob] = tX.cpu().numpy()
            t1 = tA.clone()
            nat.call("mnpbem_getri_blockinv_batched", nat.DTYPE_C64X, n, nat.ptr(t1), n, n * n, nat.ptr(piv), n,
                     nat.ptr(D), sD, 128, nat.ptr(W), n, n * n, nat.ptr(Q), n, B, st)
            inv[knob] = t1.cpu().numpy()
    finally:
        lib.mnpbem_trsm_left_rhs(prev)
        lib.mnpbem_trsm_leaf_fused(prev_f)
    assert np.all(info.cpu().numpy() == 0)
    # the folded leaf runs the stand-alone leaf's splits and MFMA order
    assert np.array_equal(out[512], out[513]) and np.array_equal(inv[512], inv[513])
    scale = np.abs(ref).max()
    e_left = np.abs(out[512] - ref).max() / scale
    e_rec = np.abs(out[0] - ref).max() / scale
    assert e_left < 3 * e_rec + 1e-5, (e_left, e_rec)
    for b in range(B):
        rinv = np.linalg.inv(A[b].astype(np.complex128))
        sc = np.abs(rinv).max()
        e1, e2 = np.abs(inv[512][b] - rinv).max() / sc, np.abs(inv[0][b] - rinv).max() / sc
        assert e1 < 3 * e2 + 1e-5, (b, e1, e2)


def _block_inverses(nat, tA, n, B, bs, st):
    """Inverse diagonal blocks of the factors in tA: 64 rows from
    mnpbem_lu_blockinv_batched, 128 rows built from those by
    mnpbem_lu_blockinv128_batched."""
    s64 = ((n + 63) // 64) * 2 * 64 * 64
    D64 = torch.empty((B, s64), dtype=torch.complex64, device=tA.device)
    nat.call("mnpbem_lu_blockinv_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(D64), s64, B, st)
    if bs == 64:
        return D64, s64
    s128 = ((n + 127) // 128) * 2 * 128 * 128
    D128 = torch.empty((B, s128), dtype=torch.complex64, device=tA.device)
    nat.call("mnpbem_lu_blockinv128_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(D64), s64,
             nat.ptr(D128), s128, B, st)
    return D128, s128


@pytest.mark.parametrize("n", [60, 128, 200, 300])
def test_lu_blockinv128_matches_numpy(cuda, n):
    """The 128-row inverse blocks against numpy's inverses of the unit-lower
    and upper triangles of each 128 x 128 diagonal block (the last one partial),
    zero outside the triangle."""
    nat = _lib()
    rng = np.random.default_rng(3 * n)
    B = 2
    A = (_rand(rng, (B, n, n), np.complex64) + (2.0 * np.sqrt(n)) * np.eye(n)[None]).astype(np.complex64)
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    st = nat.stream_handle(cuda)
    print("blockinv128", n, flush=True)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C64, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    D, sD = _block_inverses(nat, tA, n, B, 128, st)
    LU = tA.cpu().numpy().astype(np.complex128)
    out = D.cpu().numpy().reshape(B, -1, 2, 128, 128)
    for b in range(B):
        for p in range((n + 127) // 128):
            i0, nb = 128 * p, min(128, n - 128 * p)
            blk = LU[b, i0:i0 + nb, i0:i0 + nb]
            U = np.triu(blk)
            L = np.tril(blk, -1) + np.eye(nb)
            for tri, M in ((0, U), (1, L)):
                ref = np.zeros((128, 128), np.complex128)
                ref[:nb, :nb] = np.linalg.inv(M)
                err = np.abs(out[b, p, tri] - ref).max() / np.abs(ref).max()
                assert err < 1e-5, (b, p, tri, err)


@pytest.mark.parametrize("bs", [64, 128])
@pytest.mark.parametrize("n", [65, 300, 1031])
def test_getri_blockinv_leaves(cuda, n, bs):
    """mnpbem_getri_blockinv_batched against numpy's inverse, no less
    accurate than mnpbem_getri_batched on the same factors."""
    nat = _lib()
    rng = np.random.default_rng(7 * n)
    B = 2
    A = _rand(rng, (B, n, n), np.complex64)
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    st = nat.stream_handle(cuda)
    print("getri_blockinv", n, flush=True)
    nat.call("mnpbem_getrf_batched", nat.DTYPE_C64X, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
    D, sD = _block_inverses(nat, tA, n, B, bs, st)
    W = torch.empty_like(tA)
    Q = torch.empty((B, n), dtype=torch.int32, device=cuda)
    t1, t2 = tA.clone(), tA.clone()
    nat.call("mnpbem_getri_blockinv_batched", nat.DTYPE_C64X, n, nat.ptr(t1), n, n * n, nat.ptr(piv), n, nat.ptr(D),
             sD, bs, nat.ptr(W), n, n * n, nat.ptr(Q), n, B, st)
    nat.call("mnpbem_getri_batched", nat.DTYPE_C64X, n, nat.ptr(t2), n, n * n, nat.ptr(piv), n, nat.ptr(W), n, n * n,
             nat.ptr(Q), n, B, st)
    o1, o2 = t1.cpu().numpy(), t2.cpu().numpy()
    for b in range(B):
        ref = np.linalg.inv(A[b].astype(np.complex128))
        sc = np.abs(ref).max()
        e1, e2 = np.abs(o1[b] - ref).max() / sc, np.abs(o2[b] - ref).max() / sc
        assert e1 < 3 * e2 + 1e-4, (b, e1, e2)


@pytest.mark.parametrize("dtype,tol,split", [(np.complex128, 1e-12, False), (np.complex64, 2e-5, False),
                                             (np.complex64, 2e-4, True)])
@pytest.mark.parametrize("n", [1, 7, 33, 64, 65, 100, 300, 1031])
def test_getrf_nopiv_equals_partial_pivoting(cuda, dtype, tol, split, n):
    """Diagonally dominant matrices (partial pivoting keeps the diagonal):
    the factorisation without interchanges reports no interchange, returns
    the identity permutation and the factors of LAPACK's zgetrf and of the
    pivoting getrf, and its getrs solves the system."""
    import scipy.linalg as sla
    nat = _lib()
    rng = np.random.default_rng(100 + n)
    B = 3
    A = (_rand(rng, (B, n, n), dtype) + (2.0 * np.sqrt(n) + 2.0) * np.eye(n)[None]).astype(dtype)
    dt = nat.DTYPE_C128 if dtype == np.complex128 else (nat.DTYPE_C64X if split else nat.DTYPE_C64)
    st = nat.stream_handle(cuda)
    out = []
    for entry in ("mnpbem_getrf_nopiv_batched", "mnpbem_getrf_batched"):
        tA = torch.from_numpy(A.copy()).to(cuda)
        piv = torch.full((B, n), -7, dtype=torch.int32, device=cuda)
        info = torch.full((B,), -7, dtype=torch.int32, device=cuda)
        flag = torch.full((B,), -7, dtype=torch.int32, device=cuda)
        print(entry, dtype.__name__, split, n, flush=True)
        if entry == "mnpbem_getrf_nopiv_batched":
            nat.call(entry, dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), nat.ptr(flag), B, st)
            assert np.all(flag.cpu().numpy() == 0)
        else:
            nat.call(entry, dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), B, st)
        assert np.all(info.cpu().numpy() == 0)
        assert np.array_equal(piv.cpu().numpy(), np.tile(np.arange(n), (B, 1)))
        out.append(tA.cpu().numpy().astype(np.complex128))
    for b in range(B):
        lu, p = sla.lu_factor(A[b].astype(np.complex128))
        assert np.array_equal(p, np.arange(n))
        scale = np.abs(lu).max()
        assert np.abs(out[0][b] - lu).max() < tol * scale, b
        assert np.abs(out[0][b] - out[1][b]).max() < tol * scale, b
    X = _rand(rng, (B, n, 5), dtype)
    tA = torch.from_numpy(A.copy()).to(cuda)
    tX = torch.from_numpy(X.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    flag = torch.empty((B,), dtype=torch.int32, device=cuda)
    nat.call("mnpbem_getrf_nopiv_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info),
             nat.ptr(flag), B, st)
    nat.call("mnpbem_getrs_batched", dt, n, 5, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(tX), 5, n * 5, B, st)
    ref = np.linalg.solve(A.astype(np.complex128), X.astype(np.complex128))
    assert np.abs(tX.cpu().numpy() - ref).max() / np.abs(ref).max() < 10 * tol


def test_getrf_nopiv_flags_interchanges_and_zero_pivots(cuda):
    """A general random matrix needs interchanges (flag 1, as LAPACK's
    pivots show); a dominant one does not (flag 0); an exactly zero pivot is
    reported in info like getrf; batches above the 512 matrices of the
    inverse-block buffer run in chunks."""
    import scipy.linalg as sla
    nat = _lib()
    rng = np.random.default_rng(5)
    st = nat.stream_handle(cuda)
    n = 150
    A = _rand(rng, (2, n, n), np.complex128)
    A[1] += 3.0 * np.sqrt(n) * np.eye(n)
    assert not np.array_equal(sla.lu_factor(A[0])[1], np.arange(n))
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((2, n), dtype=torch.int32, device=cuda)
    info = torch.empty((2,), dtype=torch.int32, device=cuda)
    flag = torch.empty((2,), dtype=torch.int32, device=cuda)
    nat.call("mnpbem_getrf_nopiv_batched", nat.DTYPE_C128, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info),
             nat.ptr(flag), 2, st)
    f = flag.cpu().numpy().tolist()
    assert f[0] >= 1 and f[1] == 0, f
    # levels: multipliers 2 (interchange, within the growth guard of 1024) and 1e4
    T = np.array([[[1.0, 0.5], [2.0, 1.0]], [[1e-4, 1.0], [1.0, 1.0]], [[4.0, 1.0], [1.0, 3.0]]], np.complex128)
    tT = torch.from_numpy(T.copy()).to(cuda)
    p3 = torch.empty((3, 2), dtype=torch.int32, device=cuda)
    i3 = torch.empty((3,), dtype=torch.int32, device=cuda)
    f3 = torch.empty((3,), dtype=torch.int32, device=cuda)
    nat.call("mnpbem_getrf_nopiv_batched", nat.DTYPE_C128, 2, nat.ptr(tT), 2, 4, nat.ptr(p3), 2, nat.ptr(i3),
             nat.ptr(f3), 3, st)
    assert f3.cpu().numpy().tolist() == [1, 2, 0]
    Z = np.eye(40, dtype=np.complex128)
    Z[17, 17] = 0.0
    tZ = torch.from_numpy(Z[None].copy()).to(cuda)
    p1 = torch.empty((1, 40), dtype=torch.int32, device=cuda)
    i1 = torch.empty((1,), dtype=torch.int32, device=cuda)
    f1 = torch.empty((1,), dtype=torch.int32, device=cuda)
    nat.call("mnpbem_getrf_nopiv_batched", nat.DTYPE_C128, 40, nat.ptr(tZ), 40, 1600, nat.ptr(p1), 40, nat.ptr(i1),
             nat.ptr(f1), 1, st)
    assert int(i1.cpu()[0]) == 18
    B, n = 520, 70
    A = (_rand(rng, (B, n, n), np.complex64) + (2.0 * np.sqrt(n) + 2.0) * np.eye(n)[None]).astype(np.complex64)
    outs = []
    for entry in ("mnpbem_getrf_nopiv_batched", "mnpbem_getrf_batched"):
        tA = torch.from_numpy(A.copy()).to(cuda)
        piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
        info = torch.empty((B,), dtype=torch.int32, device=cuda)
        flag = torch.empty((B,), dtype=torch.int32, device=cuda)
        extra = (nat.ptr(flag),) if entry == "mnpbem_getrf_nopiv_batched" else ()
        nat.call(entry, nat.DTYPE_C64, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info), *extra, B, st)
        if extra:
            assert np.all(flag.cpu().numpy() == 0)
        outs.append(tA.cpu().numpy())
    assert np.abs(outs[0] - outs[1]).max() < 2e-5 * np.abs(outs[1]).max()


def test_factor_interchanges_fallback(cuda, monkeypatch):
    """BEMRet.factor runs G1, G2 and Delta without interchanges; when the
    factorisation records one it factors again with partial pivoting.  The
    record is forced from the host here (the flag words set to 1 after the
    call) and the solution must equal the one of the plain pivoting path."""
    import mnpbem_amd as m
    from mnpbem_amd import _native as nat
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(10.0, 9.065, 0.0708)], [m.trisphere(144, 20.0)], [[2, 1]], 1)
    pw = m.PlaneWaveRet([[1, 0, 0]], [[0, 0, 1]])
    lam = [520.0, 610.0]
    bem = m.BEMRet(cp, precision="double")
    ref = bem.solve_sweep(lam, pw)
    F = bem.factor(lam)
    assert F.npflag is not None and int(F.npflag.max()) == 0
    Fp = bem.factor(lam, interchanges=True)
    assert Fp.npflag is None
    ones = torch.ones((2,), dtype=torch.int32, device=cuda)
    orig = nat.call
    seen = []

    def call(name, *args):
        rc = orig(name, *args)
        if name == "mnpbem_getrf_nopiv_batched":
            seen.append(name)
            # flag pointer (argument 8): copy the bits of two int32 ones over the first two words
            orig("mnpbem_convert", nat.DTYPE_C128, 1, nat.ptr(ones), args[8], args[-1])
        return rc

    monkeypatch.setattr(nat, "call", call)
    F2 = bem.factor(lam)
    monkeypatch.setattr(nat, "call", orig)
    assert seen and F2.npflag is None  # factored again with interchanges
    lams = np.asarray(lam)
    X = bem.solve_batch(F2, pw.rhs(cp, lams, bem.g))
    out = pw.extinction_batch(cp, lams, X, bem.g).cpu().numpy()
    assert np.abs(out / ref - 1).max() < 1e-12


@pytest.mark.parametrize("mfma", [0, 1])
@pytest.mark.parametrize("n,C", [(2884, 8), (1030, 6), (130, 3), (512, 16), (257, 8), (66, 1), (98, 5)])
def test_gemv_t_paired_rows_and_matrix_cores(cuda, n, C, mfma):
    """mnpbem_gemv_t_batched (Y = beta Y + alpha M^T X, M the row-major
    buffer) in complex64 against numpy.  With n even and a 16-byte aligned
    buffer the product runs on the FP32 matrix cores for C <= 8 (knob
    mnpbem_gemv_mfma 1) or with two output rows per lane (knob 0); otherwise
    one row per lane.  The two-row kernel gives the same bits as the one-row
    kernel (same buffer at a misaligned address); the matrix-core kernel sums
    in another order and must agree with it to FP32 rounding."""
    nat = _lib()
    rng = np.random.default_rng(n + C)
    B = 3
    M = _rand(rng, (B, n, n), np.complex64)
    X = _rand(rng, (B, n, C), np.complex64)
    Y0 = _rand(rng, (B, n, C), np.complex64)
    al, beta = 0.7 - 0.2j, 0.5
    ref = beta * Y0.astype(np.complex128) + al * np.einsum("bji,bjc->bic", M.astype(np.complex128),
                                                            X.astype(np.complex128))
    st = nat.stream_handle(cuda)
    print("gemv_t", n, C, "mfma", mfma, flush=True)
    outs = []
    lib = nat.load()
    prev = lib.mnpbem_gemv_mfma(mfma)
    try:
        for off in (0, 1):  # complex elements before the buffer: 0 aligned, 1 misaligned (8 bytes)
            buf = torch.zeros(B * n * n + 2, dtype=torch.complex64, device=cuda)
            buf[off:off + B * n * n] = torch.from_numpy(M.reshape(-1)).to(cuda)
            tX = torch.from_numpy(X.copy()).to(cuda)
            tY = torch.from_numpy(Y0.copy()).to(cuda)
            nat.call("mnpbem_gemv_t_batched", nat.DTYPE_C64, n, C, nat.ptr_off(buf, off), n, n * n, nat.ptr(tX),
                     C, n * C, nat.ptr(tY), C, n * C, al.real, al.imag, beta, B, st)
            outs.append(tY.cpu().numpy())
    finally:
        lib.mnpbem_gemv_mfma(prev)
    scale = np.abs(ref).max()
    for o in outs:
        err = np.abs(o - ref).max() / scale
        assert err < 1e-5, err
    if mfma and C <= 8 and n % 2 == 0:
        assert np.abs(outs[0] - outs[1]).max() / scale < 1e-5
    else:
        assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("dt_name,bs", [("c64x", 128), ("c64x", 64), ("c64", 0), ("c128", 0)])
@pytest.mark.parametrize("n", [65, 300, 1031])
def test_getri_nopiv_leaves_inverse_in_w(cuda, dt_name, bs, n):
    """mnpbem_getri_nopiv_blockinv_batched on factors without interchanges:
    the inverse is left in W, A (the factors) is not written, and W holds the
    same bits that mnpbem_getri_blockinv_batched writes back into A for the
    identity permutation; both against numpy's inverse."""
    nat = _lib()
    dt = {"c64x": nat.DTYPE_C64X, "c64": nat.DTYPE_C64, "c128": nat.DTYPE_C128}[dt_name]
    npdt = np.complex128 if dt_name == "c128" else np.complex64
    rng = np.random.default_rng(11 * n + bs)
    B = 2
    A = (_rand(rng, (B, n, n), npdt) + (2.0 * np.sqrt(n) + 2.0) * np.eye(n)[None]).astype(npdt)
    tA = torch.from_numpy(A.copy()).to(cuda)
    piv = torch.empty((B, n), dtype=torch.int32, device=cuda)
    info = torch.empty((B,), dtype=torch.int32, device=cuda)
    flag = torch.zeros((B,), dtype=torch.int32, device=cuda)
    st = nat.stream_handle(cuda)
    print("getri_nopiv", dt_name, bs, n, flush=True)
    nat.call("mnpbem_getrf_nopiv_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(piv), n, nat.ptr(info),
             nat.ptr(flag), B, st)
    assert np.all(flag.cpu().numpy() == 0) and np.array_equal(piv.cpu().numpy(), np.tile(np.arange(n), (B, 1)))
    D, sD = (None, 0) if bs == 0 else _block_inverses(nat, tA, n, B, bs, st)
    LU = tA.clone()
    W = torch.full_like(tA, complex(np.nan, 0.0))
    nat.call("mnpbem_getri_nopiv_blockinv_batched", dt, n, nat.ptr(tA), n, n * n, nat.ptr(D), sD, max(bs, 64),
             nat.ptr(W), n, n * n, B, st)
    W2 = torch.empty_like(tA)
    Q = torch.empty((B, n), dtype=torch.int32, device=cuda)
    t2 = tA.clone()
    nat.call("mnpbem_getri_blockinv_batched", dt, n, nat.ptr(t2), n, n * n, nat.ptr(piv), n, nat.ptr(D), sD,
             max(bs, 64), nat.ptr(W2), n, n * n, nat.ptr(Q), n, B, st)
    assert torch.equal(tA, LU)
    w, o2 = W.cpu().numpy(), t2.cpu().numpy()
    assert np.array_equal(w, o2)
    tol = 1e-12 if dt_name == "c128" else (1e-4 if dt_name == "c64x" else 1e-5)
    for b in range(B):
        ref = np.linalg.inv(A[b].astype(np.complex128))
        assert np.abs(w[b] - ref).max() / np.abs(ref).max() < tol
