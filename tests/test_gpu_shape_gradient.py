"""GPU checks of the extinction gradient with respect to the mesh (adjoint
solve + HIP derivative kernels + torch chain rule to the vertices).

1. Per-face inputs: the kernels' derivative with respect to the normal, the
   two tangents and the area of a face, and along a rigid move of the face
   (centroid and nodes together), against central differences of the GPU
   complex128 spectrum with that one input moved (step 1e-6 / 1e-5 nm), on a
   flat rod of 216 faces.
2. Vertices: d ext / d(vertex xyz) against central differences with step
   1e-4 nm on random corner vertices of the flat rod (eps = 1.77 outside,
   polarisation along the axis; rebuilt mesh, so midpoints, centroids,
   normals, areas and quadrature nodes all move) and on random nodes of the
   curved 284-face sphere.  The truncation error of the differences,
   (h / size)^2, stays below 1e-8, their round-off below 1e-8 of the
   derivative; the bar is 1e-6 of the largest derivative component.
3. The mixed path's gradient vs the complex128 one, and the autograd wrapper.
4. A dimer of two curved spheres 1 nm apart (refined pairs across the gap,
   one closed-surface correction per sphere): node gradients of both.
"""

import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DRUDE = (10.0, 9.065, 0.0708)
EPS_OUT = 1.77


def _rod_cp(m, verts=None):
    part = m.trirod(10.0, 24.0, [10, 5, 6], triangles=True)
    if verts is not None:
        part = m.geometry.Particle(verts, part.faces, interp="flat")
    return m.ComParticle([m.EpsConst(EPS_OUT), m.EpsDrude(*DRUDE)], [part], [[2, 1]], 1)


def _spectrum(m, cp, lams, pw, green=None):
    return m.BEMRet(cp, green=green).solve_sweep(lams, pw)


def test_shape_gradient_per_face_inputs(cuda):
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_shape_gradient
    cp = _rod_cp(m)
    lams = np.array([650.0, 720.0])
    pw = m.PlaneWaveRet([[0, 0, 1]], [[1, 0, 0]])
    print("per-face gradient, rod", cp.n, flush=True)
    ext, _, Gr = extinction_with_shape_gradient(m.BEMRet(cp), lams, pw, per_face=True)
    # the functional's value: ext from the same solve
    assert np.abs(ext / _spectrum(m, cp, lams, pw) - 1).max() < 1e-12
    groups = {"nvec": range(3, 6), "tvec1": range(6, 9), "tvec2": range(9, 12), "area": range(12, 13),
              "move": range(0, 3)}
    errs = {}
    for f in (0, 61, 140, 203):
        for name, slots in groups.items():
            if name == "move":
                # rigid move of the face (centroid and its six nodes together):
                # the derivatives along pos or one node alone are those of the
                # polar self rule around a centroid 5e-7 element sizes from its
                # first nodes, beyond the reach of differences with step 1e-5
                kern = Gr[:, :, :, 0:3] + Gr[:, :, :, 13:31].reshape(Gr.shape[:3] + (6, 3)).sum(axis=3)
            else:
                kern = Gr[:, :, :, list(slots)]
            scale = np.abs(kern).max()
            for q, s in enumerate(slots):
                h = {"move": 1e-5, "area": 1e-6 * cp.area[f]}.get(name, 1e-6)
                vals = []
                for sg in (1.0, -1.0):
                    cp2 = copy.copy(cp)
                    if name == "move":
                        a = cp.pos.copy()
                        a[f, q] += sg * h
                        b = cp.nodes.copy()
                        b[f, :, q] += sg * h
                        cp2.pos, cp2.nodes = a, b
                    else:
                        a = getattr(cp, name).copy()
                        if name == "area":
                            a[f] += sg * h
                        else:
                            a[f, q] += sg * h
                        setattr(cp2, name, a)
                    vals.append(_spectrum(m, cp2, lams, pw))
                fd = (vals[0] - vals[1]) / (2 * h)
                got = kern[:, :, f, q]
                err = np.abs(got - fd).max() / scale
                errs[(f, name, q)] = err
                print("face %3d %-6s %d  kernel %+.8e  fd %+.8e  err/scale %.2e"
                      % (f, name, q, got[0, 0], fd[0, 0], err), flush=True)
    bad = {k: v for k, v in errs.items() if v > 1e-6}
    assert not bad, bad


def test_shape_gradient_vertices_fd(cuda):
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_shape_gradient
    lams = np.array([620.0, 700.0, 780.0])
    pw = m.PlaneWaveRet([[0, 0, 1]], [[1, 0, 0]])
    cp = _rod_cp(m)
    part = cp.p[0]
    print("vertex gradient, rod", cp.n, flush=True)
    ext, grads = extinction_with_shape_gradient(m.BEMRet(cp), lams, pw)
    gv = grads[0]["verts"]
    scale = np.abs(gv).max()
    rng = np.random.default_rng(7)
    h = 1e-4
    for vi in rng.choice(len(part.verts), 6, replace=False):
        for c in range(3):
            vals = []
            for sg in (1.0, -1.0):
                v = part.verts.copy()
                v[vi, c] += sg * h
                vals.append(_spectrum(m, _rod_cp(m, v), lams, pw))
            fd = (vals[0] - vals[1]) / (2 * h)
            print("rod vertex %3d %d kernel %+.8e fd %+.8e" % (vi, c, gv[0, 0, vi, c], fd[0, 0]), flush=True)
            assert np.abs(gv[:, :, vi, c] - fd).max() < 1e-6 * scale, (vi, c)
    # the FP64 CPU oracle's differences on two of them (the oracle's own flat
    # mesh built from the moved corners)
    import oracle as o
    for vi in (int(rng.integers(len(part.verts))), 40):
        for c in range(3):
            vals = []
            for sg in (1.0, -1.0):
                v = part.verts.copy()
                v[vi, c] += sg * h
                nodes, tri6 = o.flat_mesh_from_corners(v[part.faces])
                comp = o.Composite([o.EpsConst(EPS_OUT), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)],
                                   [[2, 1]])
                vals.append(o.solve_spectrum(comp, lams, [[0, 0, 1]], [[1, 0, 0]]))
            fd = (vals[0] - vals[1]) / (2 * h)
            print("rod vertex %3d %d kernel %+.8e oracle fd %+.8e" % (vi, c, gv[0, 0, vi, c], fd[0, 0]), flush=True)
            assert np.abs(gv[:, :, vi, c] - fd).max() < 1e-6 * scale, (vi, c)
    # curved sphere: gradient with respect to the 6-node faces' nodes
    pol, dr = [[1, 0, 0]], [[0, 0, 1]]
    pw2 = m.PlaneWaveRet(pol, dr)
    sph = m.trisphere(144, 20.0)
    mk = lambda nodes: m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)],
                                     [m.geometry.Particle(nodes, sph.faces6, interp="curv")], [[2, 1]], 1)
    cps = mk(sph.verts2)
    print("vertex gradient, sphere", cps.n, flush=True)
    lams2 = np.array([500.0, 530.0])
    _, gs = extinction_with_shape_gradient(m.BEMRet(cps), lams2, pw2)
    gn = gs[0]["nodes"]
    scale = np.abs(gn).max()
    for vi in rng.choice(len(sph.verts2), 4, replace=False):
        for c in range(3):
            vals = []
            for sg in (1.0, -1.0):
                v = sph.verts2.copy()
                v[vi, c] += sg * h
                vals.append(_spectrum(m, mk(v), lams2, pw2))
            fd = (vals[0] - vals[1]) / (2 * h)
            print("sphere node %3d %d kernel %+.8e fd %+.8e" % (vi, c, gn[0, 0, vi, c], fd[0, 0]), flush=True)
            assert np.abs(gn[:, :, vi, c] - fd).max() < 1e-6 * scale, (vi, c)


def test_shape_gradient_mixed_and_autograd(cuda):
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import ShapeExtinction, extinction_with_shape_gradient
    lams = np.array([640.0, 700.0])
    pw = m.PlaneWaveRet([[0, 0, 1], [0, 1, 0]], [[1, 0, 0], [1, 0, 0]])
    cp = _rod_cp(m)
    print("shape gradient mixed vs double", flush=True)
    e64, g64 = extinction_with_shape_gradient(m.BEMRet(cp), lams, pw)
    emx, gmx = extinction_with_shape_gradient(m.BEMRet(cp, precision="mixed"), lams, pw)
    assert np.abs(emx / e64 - 1).max() < 1e-10
    assert np.abs(gmx[0]["verts"] - g64[0]["verts"]).max() < 1e-8 * np.abs(g64[0]["verts"]).max()
    verts = torch.tensor(cp.p[0].verts, dtype=torch.float64, requires_grad=True)
    build = lambda v: _rod_cp(m, v)
    ext = ShapeExtinction.apply(verts, build, lams, pw, "double")
    w = torch.tensor([[1.0, 0.5], [0.25, 2.0]], dtype=torch.float64)
    (ext * w).sum().backward()
    want = (w.numpy()[:, :, None, None] * g64[0]["verts"]).sum(axis=(0, 1))
    assert np.abs(verts.grad.numpy() - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_shape_gradient_inverse_vs_factors(cuda, precision):
    """Shape gradients with the solves on explicit inverses (built in) and on
    the LU factors (inverse=False), and the Drude-parameter gradient on the
    factors (built in) and on the inverses: the same values to the
    refinement's accuracy."""
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_drude_gradient, extinction_with_shape_gradient
    lams = np.array([640.0, 700.0])
    pw = m.PlaneWaveRet([[0, 0, 1], [0, 1, 0]], [[1, 0, 0], [1, 0, 0]])
    cp = _rod_cp(m)
    print("shape gradient inverse vs factors", precision, flush=True)
    bem = m.BEMRet(cp, precision=precision)
    e1, g1 = extinction_with_shape_gradient(bem, lams, pw, inverse=False)
    e2, g2 = extinction_with_shape_gradient(bem, lams, pw)
    assert np.abs(e2 / e1 - 1).max() < 1e-10
    for key in g1[0]:
        assert np.abs(g2[0][key] - g1[0][key]).max() < 1e-8 * np.abs(g1[0][key]).max(), key
    d1 = extinction_with_drude_gradient(bem, lams, pw)
    d2 = extinction_with_drude_gradient(bem, lams, pw, inverse=True)
    for a, b in zip(d1, d2):
        assert np.abs(b - a).max() < 1e-8 * np.abs(a).max()


def test_shape_gradient_two_particles(cuda):
    """Dimer of two curved 60-face spheres 1 nm apart (refined pairs across
    the gap, one closed-surface correction per sphere): node gradients of
    both particles against central differences of the complex128 spectrum."""
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_shape_gradient
    s1 = m.trisphere(32, 8.0)
    s2 = s1.shift([9.0, 0.0, 0.0])
    pw = m.PlaneWaveRet([[1, 0, 0]], [[0, 0, 1]])
    lams = np.array([520.0, 560.0])

    def mk(n1, n2):
        parts = [m.geometry.Particle(n1, s1.faces6, interp="curv"), m.geometry.Particle(n2, s2.faces6, interp="curv")]
        return m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], parts, [[2, 1], [2, 1]], 1, 2)

    cp = mk(s1.verts2, s2.verts2)
    print("shape gradient dimer", cp.n, flush=True)
    _, grads = extinction_with_shape_gradient(m.BEMRet(cp), lams, pw)
    scale = max(np.abs(g["nodes"]).max() for g in grads)
    h = 1e-4
    rng = np.random.default_rng(3)
    for k, base in enumerate((s1.verts2, s2.verts2)):
        for vi in rng.choice(len(base), 2, replace=False):
            for c in range(3):
                vals = []
                for sg in (1.0, -1.0):
                    v = base.copy()
                    v[vi, c] += sg * h
                    cp2 = mk(v, s2.verts2) if k == 0 else mk(s1.verts2, v)
                    vals.append(_spectrum(m, cp2, lams, pw))
                fd = (vals[0] - vals[1]) / (2 * h)
                got = grads[k]["nodes"][:, :, vi, c]
                print("particle %d node %3d %d kernel %+.8e fd %+.8e" % (k, vi, c, got[0, 0], fd[0, 0]), flush=True)
                assert np.abs(got - fd).max() < 1e-6 * scale, (k, vi, c)


def test_shape_gradient_bitwise_reproducible(cuda):
    """The per-face derivative sums are formed in a fixed order (no atomics):
    two gradient computations of the same geometry give the same bits, in
    FP64 and in the mixed path."""
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import extinction_with_shape_gradient
    cp = _rod_cp(m)
    lams = np.array([650.0, 720.0])
    pw = m.PlaneWaveRet([[0, 0, 1]], [[1, 0, 0]])
    print("gradient reproducibility, rod", cp.n, flush=True)
    for precision in ("double", "mixed"):
        bem = m.BEMRet(cp, precision=precision)
        e1, g1, r1 = extinction_with_shape_gradient(bem, lams, pw, per_face=True)
        e2, g2, r2 = extinction_with_shape_gradient(bem, lams, pw, per_face=True)
        assert np.array_equal(e1, e2) and np.array_equal(r1, r2), precision
        for d1, d2 in zip(np.atleast_1d(g1), np.atleast_1d(g2)):  # per-particle dicts of vertex gradients
            assert d1.keys() == d2.keys() and all(np.array_equal(d1[k], d2[k]) for k in d1), precision
