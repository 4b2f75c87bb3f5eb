"""GPU parity of BEMRetIter (GMRES on the GPU) and of the H-matrix kernels
(ACA, matvec, densify) through the C ABI.

Oracles: the FP64 CPU direct solve (oracle.bem_oracle), the restated scipy
GMRES path (oracle.iter_oracle) and the CPU ACA / H-matrix restatement
(oracle.hmatrix_oracle), fed the same matrices.  Tolerances are written at
each assertion; the iterative solves are run at tol 1e-10 so that their
agreement with the direct solve is limited by the GMRES tolerance times the
condition of the system, not by the kernels.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle as o  # noqa: E402
from oracle.hmatrix_oracle import OracleClusterTree, OracleHMatrix  # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)
POL2 = ([[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]])


@pytest.fixture(scope="module")
def sphere284(cuda):
    import mnpbem_amd as m
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(144, 20.0)], [[2, 1]], 1, interp="curv")
    nodes, tri6 = o.trisphere_mesh(144, 20.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    return cp, comp, o.GreenSetup(comp)


def _rel(a, b):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_iter_dense_matches_direct_oracle(sphere284, precision):
    import mnpbem_amd as m
    cp, comp, ref = sphere284
    print("BEMRetIter dense", precision, flush=True)
    # GMRES at 1e-11: the distance to the direct solution is bounded by the
    # tolerance times the condition of the system (~100-1000 here)
    bem = m.BEMRetIter(cp, precision=precision, tol=1e-11, maxit=50)
    exc = m.PlaneWaveRet(*POL2)
    sig, _ = bem.solve(exc.potential(cp, 550.0))
    d = o.BEMRetOracle(ref).solve(o.planewave_potential(comp, *POL2, 550.0))
    for f in ("sig1", "sig2", "h1", "h2"):
        assert _rel(sig.get(f), d[f]) < 1e-8, f
    flag, relres, it = bem.info()
    assert flag[-1] == 0 and relres[-1] < 1e-11
    # the direct preconditioner makes GMRES converge in a few iterations
    assert it[-1][1] <= (2 if precision == "double" else 6)


def test_iter_multipol_matches_sequential(sphere284):
    """tests/test_iter_multipol.py of the reference: a two-polarisation solve
    equals the two single-polarisation solves."""
    import mnpbem_amd as m
    cp, _, _ = sphere284
    print("BEMRetIter multipol", flush=True)
    bm = m.BEMRetIter(cp, tol=1e-12)
    sm, _ = bm.solve(m.PlaneWaveRet(*POL2).potential(cp, 600.0))
    for q in range(2):
        b1 = m.BEMRetIter(cp, tol=1e-12)
        s1, _ = b1.solve(m.PlaneWaveRet([POL2[0][q]], [POL2[1][q]]).potential(cp, 600.0))
        for f in ("sig1", "sig2"):
            assert _rel(sm.get(f)[:, q], s1.get(f).cpu().numpy()) < 1e-10
        for f in ("h1", "h2"):
            assert _rel(sm.get(f)[:, :, q], s1.get(f).cpu().numpy()) < 1e-10


def test_iter_unpreconditioned_and_precond_only(cuda):
    import mnpbem_amd as m
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(32, 10.0)], [[2, 1]], 1)
    nodes, tri6 = o.trisphere_mesh(32, 10.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    ref = o.GreenSetup(comp)
    print("BEMRetIter precond None / maxit 0, N=60", flush=True)
    exc = m.PlaneWaveRet([[1, 0, 0]], [[0, 0, 1]]).potential(cp, 520.0)
    d = o.BEMRetOracle(ref).solve(o.planewave_potential(comp, [[1, 0, 0]], [[0, 0, 1]], 520.0))
    # full GMRES (restart = 8N) converges within 8N iterations
    b0 = m.BEMRetIter(cp, precond=None, tol=1e-9, maxit=2, restart=8 * cp.n)
    s0, _ = b0.solve(exc)
    flag, relres, it = b0.info()
    assert flag[-1] == 0 and relres[-1] < 1e-9
    assert _rel(s0.sig2, d["sig2"][:, 0]) < 1e-5
    bp = m.BEMRetIter(cp, maxit=0)
    sp, _ = bp.solve(exc)
    assert _rel(sp.sig2, d["sig2"][:, 0]) < 1e-10


def test_iter_sweep_matches_direct_sweep(sphere284):
    import mnpbem_amd as m
    cp, _, _ = sphere284
    print("BEMRetIter sweep", flush=True)
    lams = np.linspace(420.0, 780.0, 5)
    exc = m.PlaneWaveRet(*POL2)
    e_it = m.BEMRetIter(cp, precision="mixed", tol=1e-11).solve_sweep(lams, exc, chunk=3)
    e_d = m.BEMRet(cp).solve_sweep(lams, exc, chunk=5)
    assert np.max(np.abs(e_it / e_d - 1)) < 1e-9


# ------------------------------------------------------------------ H-matrices
def _points(n, seed=42):
    rng = np.random.RandomState(seed)
    ph = rng.uniform(0, 2 * np.pi, n)
    th = np.arccos(rng.uniform(-1, 1, n))
    return 10.0 * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)


def _kernel(pos):
    def fun(r, c):
        d = np.linalg.norm(pos[r] - pos[c], axis=1)
        return 1.0 / (4.0 * np.pi * np.maximum(d, 1e-10))
    return fun


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-5), ("complex64", 1e-5)])
def test_hmatrix_kernel_1_over_r(cuda, dtype, tol):
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    print("hmatrix 1/r", dtype, flush=True)
    pos = _points(800)
    fun = _kernel(pos)
    n = len(pos)
    rr, cc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    A = fun(rr.ravel(), cc.ravel()).reshape(n, n)
    t = ClusterTree(pos, cleaf=16)
    h = HMatrix.from_func(t, fun, htol=1e-6, kmax=64, dtype=dtype)
    ho = OracleHMatrix(OracleClusterTree(pos, cleaf=16), htol=1e-6, kmax=64).aca(fun)
    assert np.linalg.norm(h.full(np) - A) / np.linalg.norm(A) < tol
    # same block partition; ranks within one of the CPU ACA's (pivot ties may differ)
    assert len(h.ranks) == len(ho.lhs)
    assert np.max(np.abs(h.ranks - ho.ranks())) <= 1
    assert abs(h.compression() - ho.compression()) < 0.02
    X = np.random.RandomState(2).randn(n, 3) + 1j * np.random.RandomState(3).randn(n, 3)
    Y = h.mtimes_vec(X)
    assert np.linalg.norm(Y - A @ X) / np.linalg.norm(A @ X) < tol
    Yf = h.full(np) @ X
    assert np.linalg.norm(Y - Yf) / np.linalg.norm(Yf) < (1e-13 if dtype == "complex128" else 1e-6)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_hmatrix_packed_product_matches_block_product(cuda, dtype):
    """The product on the packed panels (matvec_device) against the product
    on the block storage (matvec_blocks_device), same stored values, both
    accumulating in complex128: 1, 4, 7 and 16 columns, row strides above
    the column count, overwrite and accumulate, leaves of up to 100 rows
    (two 64-row groups each)."""
    from mnpbem_amd import _native as nat
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    print("hmatrix packed vs block product", dtype, flush=True)
    pos = _points(3000, seed=4)
    t = ClusterTree(pos, cleaf=100)
    h = HMatrix.from_func(t, _kernel(pos), htol=1e-6, kmax=64, dtype=dtype)
    assert (h.ranks > 0).any() and t.n == 3000
    n, ld = t.n, 17
    g = torch.Generator(device="cpu").manual_seed(11)
    for ncol in (1, 4, 7, 16):
        X = torch.randn(n, ld, dtype=torch.complex128, generator=g).cuda()
        Y0 = torch.randn(n, ld, dtype=torch.complex128, generator=g).cuda()
        for acc in (False, True):
            Y1, Y2 = Y0.clone(), Y0.clone()
            h.matvec_device(nat.ptr(X), ld, nat.ptr(Y1), ld, ncol, accumulate=acc)
            h.matvec_blocks_device(nat.ptr(X), ld, nat.ptr(Y2), ld, ncol, accumulate=acc)
            torch.cuda.synchronize()
            d = (Y1[:, :ncol] - Y2[:, :ncol]).abs().max().item() / Y2[:, :ncol].abs().max().item()
            assert d < 1e-13, (ncol, acc, d)
            assert torch.equal(Y1[:, ncol:], Y0[:, ncol:])  # columns beyond ncol untouched


def test_hmatrix_release_blocks_keeps_products(cuda):
    """After release_blocks the products on the packed panels are unchanged
    bit for bit, the host tables still answer, and a method that reads the
    freed block storage raises instead of reading freed memory."""
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    print("hmatrix release_blocks", flush=True)
    pos = _points(900, seed=6)
    t = ClusterTree(pos, cleaf=16)
    h = HMatrix.from_func(t, _kernel(pos), htol=1e-6, kmax=64, dtype="complex64")
    X = np.random.RandomState(3).randn(len(pos), 3) + 1j * np.random.RandomState(4).randn(len(pos), 3)
    Y0 = h.mtimes_vec(X)
    st0 = h.stat()
    assert h.release_blocks() is h
    assert np.array_equal(h.mtimes_vec(X), Y0)
    assert h.stat() == st0
    for call in (lambda: h.full(np), lambda: h.truncate(1e-4), lambda: h.diag(np)):
        with pytest.raises(RuntimeError, match="released"):
            call()


def test_hmatrix_packed_product_follows_block_changes(cuda):
    """The packed panels are rebuilt after the blocks change in place
    (fillval, truncate, scaling): each product equals the densified matrix's."""
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    print("hmatrix packed product after changes", flush=True)
    pos = _points(900, seed=5)
    fun = _kernel(pos)
    n = len(pos)
    t = ClusterTree(pos, cleaf=16)
    h = HMatrix.from_func(t, fun, htol=1e-8, kmax=64)
    X = np.random.RandomState(7).randn(n, 2) + 1j * np.random.RandomState(8).randn(n, 2)

    def check(hm):
        Y = hm.mtimes_vec(X)
        Yf = hm.full(np) @ X
        assert np.linalg.norm(Y - Yf) / np.linalg.norm(Yf) < 1e-13

    check(h)
    h.fillval(lambda r, c: 3.0 * fun(r, c))
    check(h)
    h.truncate(1e-4)
    check(h)
    check(2.5 * h)
    check(h + h)


def test_hmatrix_two_particles_known_rank(cuda):
    """Two well-separated particles and a separable kernel: every admissible
    block has exact rank 2, ACA must stop at rank <= 3 with round-off error."""
    from mnpbem_amd.hmatrix import ClusterTree, HMatrix
    print("hmatrix separable", flush=True)
    p1, p2 = _points(300), _points(200, seed=9)
    p2[:, 0] += 100.0
    pos = np.concatenate([p1, p2])
    ip = np.repeat([0, 1], [300, 200])
    f1 = np.exp(1j * pos[:, 0] / 7.0)
    f2 = pos[:, 1] + 2.0
    g1, g2 = pos[:, 2] ** 2 + 1.0, np.cos(pos[:, 0])
    fun = lambda r, c: f1[r] * g1[c] + f2[r] * g2[c]
    t = ClusterTree(pos, cleaf=20, ipart_arr=ip)
    h = HMatrix.from_func(t, fun, htol=1e-12, kmax=30)
    n = len(pos)
    rr, cc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    A = fun(rr.ravel(), cc.ravel()).reshape(n, n)
    assert h.ranks.max() <= 3
    assert np.linalg.norm(h.full(np) - A) / np.linalg.norm(A) < 1e-12


@pytest.mark.parametrize("key", ["G", "H1"])
def test_aca_green_matches_cpu_aca(sphere284, key):
    """ACA of the assembled Green matrix on the GPU vs the CPU ACA of the
    same matrix: same partition, ranks within one, both within htol."""
    from mnpbem_amd.hmatrix import ACACompGreenRet
    from oracle.hmatrix_oracle import kaware_fadmiss
    cp, _, _ = sphere284
    print("aca green", key, flush=True)
    aca = ACACompGreenRet(cp, htol=1e-6, kmax=50, cleaf=16)
    lam = 600.0
    h = aca.eval(0, 0, key, lam)
    M = aca.g.eval(0, 0, key, lam).cpu().numpy()
    ho = OracleHMatrix(OracleClusterTree(cp.pos, cleaf=16), htol=1e-6, kmax=50,
                       fadmiss=kaware_fadmiss(2 * np.pi / lam)).aca(lambda r, c: M[r, c])
    assert len(h.ranks) == len(ho.lhs) and np.max(np.abs(h.ranks - ho.ranks())) <= 1
    err = np.linalg.norm(h.full(np) - M) / np.linalg.norm(M)
    erro = np.linalg.norm(ho.full() - M) / np.linalg.norm(M)
    assert err < 1e-5 and erro < 1e-5, (err, erro)


@pytest.mark.parametrize("hdt", ["complex128", "complex64"])
def test_iter_hmatrix_matches_direct(cuda, hdt):
    """BEMRetIter(hmatrix=True): GMRES over the ACA operator; its solution
    differs from the dense direct one by the compression error (htol)."""
    import mnpbem_amd as m
    print("BEMRetIter hmatrix", hdt, flush=True)
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(400, 40.0)], [[2, 1]], 1)
    lams = np.array([500.0, 650.0])
    exc = m.PlaneWaveRet(*POL2)
    bi = m.BEMRetIter(cp, hmatrix=True, htol=1e-8, kmax=60, cleaf=32, tol=1e-10, maxit=30, hmat_dtype=hdt)
    e_h = bi.solve_sweep(lams, exc, chunk=2)
    e_d = m.BEMRet(cp).solve_sweep(lams, exc)
    lim = 1e-6 if hdt == "complex128" else 2e-5
    assert np.max(np.abs(e_h / e_d - 1)) < lim
    assert all(f == 0 for f in bi.info()[0])
    # at N = 796 with htol 1e-8 and the wavelength-aware admissibility the
    # factors cost about as much as the dense blocks they replace
    c = bi.hinfo()
    assert 0 < c < 1.5


@pytest.mark.parametrize("precision,hmatrix", [("double", False), ("mixed", False), ("mixed", True)])
def test_iter_block_preconditioner_dimer(cuda, precision, hmatrix):
    """preconditioner='block' (each particle's direct solve on its own faces):
    on a 1-nm-gap dimer of two 284-face spheres GMRES needs more than one
    iteration (the coupling is not in the preconditioner) and converges to
    the dense direct solution."""
    import mnpbem_amd as m
    print("BEMRetIter block preconditioner", precision, hmatrix, flush=True)
    p1 = m.trisphere(144, 10.0)
    p2 = p1.shift([11.0, 0.0, 0.0])
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2)
    lams = np.array([520.0, 600.0])
    exc = m.PlaneWaveRet(*POL2)
    opts = dict(hmatrix=True, htol=1e-10, kmax=80, cleaf=32) if hmatrix else {}
    bi = m.BEMRetIter(cp, precision=precision, preconditioner="block", tol=1e-10, maxit=60, restart=60, **opts)
    e_b = bi.solve_sweep(lams, exc, chunk=2)
    e_d = m.BEMRet(cp).solve_sweep(lams, exc)
    flags, rel, its = bi.info()
    assert all(f == 0 for f in flags)
    assert max(int(i[1]) for i in its) > 1
    assert np.max(np.abs(e_b / e_d - 1)) < (1e-8 if not hmatrix else 1e-6)


@pytest.mark.parametrize("dimer", [False, True])
def test_aca_generated_in_place_matches_assembled(sphere284, dimer):
    """ACACompGreenRet.eval generates every entry in the kernels (dense leaf
    blocks and ACA rows / columns, no n x n matrix): its dense blocks equal
    the assembled matrix's, the ranks agree within one with the ACA of the
    assembled matrix, and both reproduce the matrix to htol (G, H1 inside;
    G, H2 outside; a one-particle sphere and a 1-nm-gap dimer whose inside
    Green function couples the two particles)."""
    import mnpbem_amd as m
    from mnpbem_amd.hmatrix import ACACompGreenRet, HMatrix
    if dimer:
        p1 = m.trisphere(144, 10.0)
        cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p1, p1.shift([11.0, 0.0, 0.0])],
                           [[2, 1], [2, 1]], 1, 2)
    else:
        cp = sphere284[0]
    lam = 580.0
    htol = 1e-8
    aca = ACACompGreenRet(cp, htol=htol, kmax=60, cleaf=16)
    for side, key in [(0, "G"), (0, "H1"), (1, "G"), (1, "H2")]:
        print("in-place ACA", "dimer" if dimer else "sphere", side, key, flush=True)
        h = aca.eval(side, side, key, lam)
        M = aca.g.eval(side, side, key, lam)
        ha = HMatrix(aca.tree, htol=htol, kmax=60, fadmiss=aca.fadmiss_for(lam)).aca(M)
        Mh = M.cpu().numpy()
        assert np.array_equal(h._blk_host[:h._nd], ha._blk_host[:ha._nd])
        assert _rel(h._val, ha._val.cpu().numpy()) < 1e-14, (side, key)
        assert np.max(np.abs(h.ranks - ha.ranks)) <= 1
        for hm in (h, ha):
            assert np.linalg.norm(hm.full(np) - Mh) / np.linalg.norm(Mh) < 10 * htol, (side, key)
    # inside and outside share no material here: no particle connects them
    assert aca.eval(0, 1, "G", lam) == 0


@pytest.mark.parametrize("precision,tol", [("double", 1e-9), ("mixed", 1e-3)])
def test_bem_solve_with_inverses_matches_factors(cuda, precision, tol):
    """mnpbem_bem_solve_inv_batched (explicit inverses of G1, G2, Sigma from
    BEMRet.invert_factors) against mnpbem_bem_solve_batched on the factors of
    the same factorisation: the same direct solve, to the inverse's accuracy
    (complex128 here to ~1e-12, the bf16x3 inverses to ~1e-5)."""
    import mnpbem_amd as m
    from mnpbem_amd import _native as nat
    print("bem solve with inverses", precision, flush=True)
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(144, 20.0)], [[2, 1]], 1)
    lams = np.array([520.0, 600.0])
    exc = m.PlaneWaveRet(*POL2)
    bem = m.BEMRet(cp, precision=precision)
    F = bem.factor(lams)
    R = exc.rhs(cp, lams, bem.g)
    P = R.shape[2] // 8
    cdt = torch.complex128 if precision == "double" else torch.complex64
    Rc = R.to(cdt)
    X1, W1 = torch.empty_like(Rc), torch.empty_like(Rc)
    bem._solve_direct(F, Rc, X1, W1, P)
    bem.invert_factors(F)
    X3 = torch.empty_like(Rc)
    bem._solve_direct(F, Rc, X3, W1, P)  # takes the inverses from here on
    X2, W2 = torch.empty_like(Rc), torch.empty_like(Rc)
    n, nn = F.n, F.n * F.n
    nat.call("mnpbem_bem_solve_inv_batched", F.dtype, n, P, nat.ptr(F.GG[0]), nat.ptr(F.GG[1]), nat.ptr(F.HH[0]),
             nat.ptr(F.Et), nat.ptr(F.HH[1]), n, nn, nat.ptr(Rc), nat.ptr(W2), nat.ptr(X2), 8 * P, n * 8 * P,
             nat.ptr(bem.g.nvec), nat.ptr(F.par), 0, F.B, nat.stream_handle(cuda))
    a, b = X1.cpu().numpy(), X2.cpu().numpy()
    assert np.abs(a - b).max() / np.abs(a).max() < tol
    assert torch.equal(X2, X3)


def test_iter_block_preconditioner_inverse_vs_factors(cuda):
    """The block preconditioner with explicit inverses (built in) and with the
    triangular factors (precond_inverse=False): the same GMRES solution of a
    dimer to the solver tolerance, within one iteration of each other."""
    import mnpbem_amd as m
    print("block preconditioner inverse vs factors", flush=True)
    p1 = m.trisphere(144, 10.0)
    p2 = p1.shift([11.0, 0.0, 0.0])
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2)
    lams = np.array([520.0, 600.0])
    exc = m.PlaneWaveRet(*POL2)
    out = {}
    for inv in (True, False):
        bi = m.BEMRetIter(cp, precision="mixed", preconditioner="block", tol=1e-10, maxit=60, restart=60,
                          precond_inverse=inv)
        out[inv] = (bi.solve_sweep(lams, exc, chunk=2), bi.info())
    assert np.max(np.abs(out[True][0] / out[False][0] - 1)) < 1e-8
    its = [max(int(i[1]) for i in out[k][1][2]) for k in (True, False)]
    assert abs(its[0] - its[1]) <= 1, its
