"""GPU parity at the benchmark workloads themselves (not reduced meshes).

* 2884-face Drude-gold sphere (trisphere(1444, 40), eps_inf 10, wp 9.065 eV,
  gamma 0.0708 eV, in vacuum; x and y polarised plane waves along +z):
  complex128 and mixed paths against the FP64 CPU oracle at 520 nm and two
  sweep wavelengths (1e-10 / 1e-9); the full 64-wavelength mixed sweep
  against the complex128 GPU sweep (1e-8, the accuracy target of the mixed
  path); the hipGraph-captured 64-wavelength step against the eager sweep.
* 2880-face gold rod (trirod(20, 60, [31, 13, 26]), flat triangles, in
  eps 1.77; polarisation along the axis, propagation along x) at its
  longitudinal resonance: the adjoint vertex gradient of the mixed path
  against central differences (step 1e-4 nm) of the complex128 spectrum on
  seeded random vertices, and the Drude-parameter gradient against central
  differences; bar 1e-6 of the gradient's scale.
* 25 000-face dimer (two trisphere_fibonacci(6252, 100) spheres, 2 nm gap):
  BEMRetIter on ACA H-matrices (leaf 64, htol 1e-6, rank <= 30, GMRES tol
  1e-6) against the dense mixed direct solve at 600 nm; the difference is
  bounded by the compression and GMRES tolerances (1e-5).
* The bf16x3 GEMM on the sweep's own large shapes (1442 x 1442 x 1442 and
  2884 x 1442 x 1442, batch 2) against complex128 products.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle as o  # noqa: E402

DRUDE = (10.0, 9.065, 0.0708)
POL2 = ([[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]])
SWEEP = np.linspace(400.0, 800.0, 64)


def _free():
    import gc
    gc.collect()
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def sphere2884(cuda):
    import mnpbem_amd as m
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(1444, 40.0)], [[2, 1]], 1,
                       interp="curv")
    assert cp.n == 2884
    return cp


@pytest.fixture(scope="module")
def sweep64_double(sphere2884):
    import mnpbem_amd as m
    print("sphere 2884: complex128 sweep, 64 wavelengths", flush=True)
    ext = m.BEMRet(sphere2884, precision="double").solve_sweep(SWEEP, m.PlaneWaveRet(*POL2), chunk=32)
    _free()
    return ext


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("mixed", 1e-9)])
def test_sphere2884_against_oracle(sphere2884, precision, tol):
    import mnpbem_amd as m
    lams = np.array([520.0, SWEEP[9], SWEEP[50]])
    print("sphere 2884 vs oracle", precision, lams, flush=True)
    bem = m.BEMRet(sphere2884, precision=precision)
    ext = bem.solve_sweep(lams, m.PlaneWaveRet(*POL2), chunk=3)
    nodes, tri6 = o.trisphere_mesh(1444, 40.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    ref = o.solve_spectrum(comp, lams, *POL2)
    err = float(np.max(np.abs(ext / ref - 1)))
    print("max rel ext err vs oracle %.2e" % err, flush=True)
    assert err < tol, err
    if precision == "mixed":
        assert np.max(bem.last_residual) <= bem.refine_tol
    _free()


def test_sphere2884_mixed_sweep_against_complex128(sphere2884, sweep64_double):
    import mnpbem_amd as m
    print("sphere 2884: mixed sweep, 64 wavelengths in one chunk", flush=True)
    bem = m.BEMRet(sphere2884, precision="mixed")
    ext = bem.solve_sweep(SWEEP, m.PlaneWaveRet(*POL2), chunk=64)
    err = float(np.max(np.abs(ext / sweep64_double - 1)))
    print("max rel ext err vs complex128 %.2e, refinement solves %d" % (err, bem.last_refine_iters), flush=True)
    assert err < 1e-8, err
    assert np.max(bem.last_residual) <= bem.refine_tol
    _free()


def test_sphere2884_graph_replay(sphere2884, sweep64_double):
    """The captured 64-wavelength step replayed twice, with eager work in
    between: each replay equals the eager sweep and meets the 1e-8 target."""
    import mnpbem_amd as m
    from mnpbem_amd.bem import SweepGraph
    print("sphere 2884: SweepGraph B=64", flush=True)
    bem = m.BEMRet(sphere2884, precision="mixed")
    pw = m.PlaneWaveRet(*POL2)
    sg = SweepGraph(bem, pw, 64, iters=3)
    first = sg.run(SWEEP)
    eager = bem.solve_sweep(SWEEP[:4], pw, chunk=4)   # eager work between replays
    second = sg.run(SWEEP)
    assert np.array_equal(first, second)
    assert np.max(np.abs(eager / sweep64_double[:4] - 1)) < 1e-8
    err = float(np.max(np.abs(second / sweep64_double - 1)))
    print("replay vs complex128 %.2e" % err, flush=True)
    assert err < 1e-8, err
    del sg
    _free()


# ---------------------------------------------------------------------------
# 2880-face rod: adjoint gradients against central differences
# ---------------------------------------------------------------------------

ROD_LAM = 720.0   # longitudinal resonance of the rod on the 600-800 nm scan


def _rod(m, verts=None, drude=DRUDE):
    part = m.trirod(20.0, 60.0, [31, 13, 26], triangles=True)
    if verts is not None:
        part = m.geometry.Particle(verts, part.faces, interp="flat")
    return m.ComParticle([m.EpsConst(1.77), m.EpsDrude(*drude)], [part], [[2, 1]], 1)


def test_rod2880_vertex_gradient_fd(cuda):
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_shape_gradient
    cp = _rod(m)
    assert cp.n == 2880
    verts = cp.p[0].verts
    pw = m.PlaneWaveRet([[0, 0, 1]], [[1, 0, 0]])
    lams = np.array([ROD_LAM])
    print("rod 2880: mixed vertex gradient at %.0f nm" % ROD_LAM, flush=True)
    ext, grads = extinction_with_shape_gradient(m.BEMRet(cp, precision="mixed"), lams, pw, chunk=1)
    gv = grads[0]["verts"][0, 0]
    scale = np.abs(gv).max()
    # the resonance: the spectrum's maximum on a coarse scan sits at 720 nm
    scan = m.BEMRet(cp, precision="mixed").solve_sweep(np.array([680.0, 700.0, 720.0, 740.0, 760.0]), pw)
    assert int(np.argmax(scan[:, 0])) == 2, scan[:, 0]
    rng = np.random.default_rng(11)
    h = 1e-4
    worst = 0.0
    for vi in rng.choice(len(verts), 4, replace=False):
        for c in range(3):
            vals = []
            for sg in (1.0, -1.0):
                v = verts.copy()
                v[vi, c] += sg * h
                vals.append(m.BEMRet(_rod(m, v)).solve_sweep(lams, pw)[0, 0])
            fd = (vals[0] - vals[1]) / (2 * h)
            e = abs(gv[vi, c] - fd) / scale
            worst = max(worst, e)
            print("vertex %4d %d  adjoint %+.10e  fd %+.10e  err/scale %.2e" % (vi, c, gv[vi, c], fd, e),
                  flush=True)
    assert worst < 1e-6, worst
    _free()


def test_rod2880_drude_gradient_fd(cuda):
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_drude_gradient
    pw = m.PlaneWaveRet([[0, 0, 1]], [[1, 0, 0]])
    lams = np.array([ROD_LAM])
    print("rod 2880: mixed Drude gradient at %.0f nm" % ROD_LAM, flush=True)
    ext, grad = extinction_with_drude_gradient(m.BEMRet(_rod(m), precision="mixed"), lams, pw, chunk=1)
    g = grad[0, 0]
    scale = np.abs(g).max()
    for q in range(3):
        h = 1e-6 * DRUDE[q]
        vals = []
        for sg in (1.0, -1.0):
            d = list(DRUDE)
            d[q] += sg * h
            vals.append(m.BEMRet(_rod(m, drude=tuple(d))).solve_sweep(lams, pw)[0, 0])
        fd = (vals[0] - vals[1]) / (2 * h)
        print("param %d adjoint %+.10e fd %+.10e" % (q, g[q], fd), flush=True)
        assert abs(g[q] - fd) < 1e-6 * scale, (q, g[q], fd)
    _free()


# ---------------------------------------------------------------------------
# 25 000-face dimer: H-matrix GMRES against the dense direct solve
# ---------------------------------------------------------------------------

def test_dimer25k_hmatrix_gmres_against_direct(cuda):
    import mnpbem_amd as m
    print("dimer 25000: building meshes", flush=True)
    p1 = m.trisphere_fibonacci(6252, 100.0)
    p2 = p1.shift([102.0, 0.0, 0.0])
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [p1, p2], [[2, 1], [2, 1]], 1, 2)
    assert cp.n == 25000
    pw = m.PlaneWaveRet([[1.0, 0.0, 0.0]], [[0.0, 0.0, 1.0]])
    lam = np.array([600.0])
    _free()
    torch.cuda.reset_peak_memory_stats()
    bi = m.BEMRetIter(cp, hmatrix=True, htol=1e-6, kmax=30, cleaf=64, tol=1e-6, restart=50, maxit=20,
                      precision="mixed", hmat_dtype="complex64")
    # several particles: the default preconditioner is the per-particle
    # direct solve, shared by the two translated spheres
    assert len(bi._blocks) == 2 and len(bi._block_bems) == 1
    print("dimer 25000: H-matrix GMRES at 600 nm", flush=True)
    e_it = bi.solve_sweep(lam, pw, chunk=1)
    flag, rel, it = bi.info()
    peak = torch.cuda.max_memory_allocated()
    seg = max(sg["total_size"] for sg in torch.cuda.memory_snapshot())
    print("GMRES flag %s relres %s it %s; peak device memory %.2f GB, largest segment %.2f GB"
          % (flag, rel, it, peak / 1e9, seg / 1e9), flush=True)
    assert all(f == 0 for f in flag) and np.max(rel) <= 1e-6
    # the H-matrix entries are generated in place and the preconditioner
    # factors 12500-face blocks: no allocation reaches one 25000 x 25000
    # complex64 matrix (5 GB; the four complex128 Green matrices the ACA
    # used to read took 40 GB), and the whole solve stays near one
    # complex128 matrix's 10 GB
    assert seg < 8 * cp.n * cp.n, seg
    assert peak < 1.25 * 16 * cp.n * cp.n, peak
    st = [h.stat() for h in bi.last_hmats[0]]
    assert all(s["rank_max"] <= 30 for s in st)
    assert all(s["compression"] < 0.3 for s in st)
    del bi
    _free()
    print("dimer 25000: dense mixed direct solve", flush=True)
    e_d = m.BEMRet(cp, precision="mixed").solve_sweep(lam, pw, chunk=1)
    err = float(np.max(np.abs(e_it / e_d - 1)))
    print("H-matrix GMRES vs direct: %.2e" % err, flush=True)
    assert err < 1e-5, err
    _free()


# ---------------------------------------------------------------------------
# bf16x3 GEMM on the sweep's large shapes
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("M,N,K", [(1442, 1442, 1442), (2884, 1442, 1442), (192, 2884, 184)])
def test_bf16x3_gemm_sweep_shapes(cuda, M, N, K):
    """C -= A B with complex64 operands on the split-bf16 kernel against the
    complex128 product of the same complex64 inputs: the split keeps ~16 bits
    of each operand, so the error is ~1e-5 of the product's scale."""
    from mnpbem_amd import _native as nat
    g = torch.Generator(device="cpu").manual_seed(M + N + K)
    B = 2
    A = torch.randn((B, M, K), dtype=torch.complex64, generator=g).to(cuda)
    Bm = torch.randn((B, K, N), dtype=torch.complex64, generator=g).to(cuda)
    C = torch.randn((B, M, N), dtype=torch.complex64, generator=g).to(cuda)
    ref = C.to(torch.complex128) - torch.matmul(A.to(torch.complex128), Bm.to(torch.complex128))
    print("bf16x3 gemm", M, N, K, flush=True)
    nat.call("mnpbem_cgemm_sub_batched", nat.DTYPE_C64X, M, N, K, nat.ptr(A), K, M * K, nat.ptr(Bm), N, K * N,
             nat.ptr(C), N, M * N, B, nat.stream_handle(cuda))
    err = float((C.to(torch.complex128) - ref).abs().max() / ref.abs().max())
    assert err < 1e-4, err
    del A, Bm, C, ref
    _free()
