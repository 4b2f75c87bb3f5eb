"""CPU pins of the golden-spiral meshes (Fibonacci sphere, ellipsoid) and of
flat triangles carried as quadratic elements, in the oracle and in the
product's host-side mesh code (no GPU).

Independent checks: the golden-spiral closed form; a triangulation of points
on a sphere is Delaunay on the sphere exactly when it is the 3-D convex
hull of the points (so the faces must equal scipy's ConvexHull facets);
closed-manifold counts (2n - 4 faces, every edge in two faces, each directed
edge once); outward normals; area and enclosed volume against pi d^2,
4/3 pi a b c and the ellipsoid's surface integral; flat-element geometry and
quadrature points against the straight-triangle formulas of particle.py.
"""

import numpy as np
import pytest
from scipy.spatial import ConvexHull

import oracle as o


@pytest.mark.parametrize("n", [60, 400, 1001])
def test_golden_spiral_closed_form(n):
    v = o.golden_spiral(n)
    k = np.arange(n) + 0.5
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
    assert np.allclose(v[:, 2], 1 - 2 * k / n, atol=1e-14)
    az = np.unwrap(np.arctan2(v[:, 1], v[:, 0]))
    d = np.mod(np.diff(az), 2 * np.pi)
    assert np.allclose(d, np.mod(np.pi * (1 + 5 ** 0.5), 2 * np.pi), atol=1e-9)


@pytest.mark.parametrize("n", [144, 400, 1444])
def test_spherical_delaunay_is_the_convex_hull(n):
    v = o.golden_spiral(n)
    f = o.spherical_delaunay(v)
    assert len(f) == 2 * n - 4
    hull = ConvexHull(v)
    assert {tuple(sorted(t)) for t in f} == {tuple(sorted(t)) for t in hull.simplices}
    directed = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    assert len({tuple(e) for e in directed}) == len(directed)          # each directed edge once
    und = np.sort(directed, axis=1)
    _, cnt = np.unique(und, axis=0, return_counts=True)
    assert np.all(cnt == 2)                                            # closed 2-manifold
    # MNPBEM orders faces clockwise seen from outside: the outward normal is
    # (v0 - v1) x (v2 - v1) (particle.py _norm_flat)
    c = v[f].mean(axis=1)
    nrm = np.cross(v[f[:, 0]] - v[f[:, 1]], v[f[:, 2]] - v[f[:, 1]])
    assert np.all(np.sum(nrm * c, axis=1) > 0)                         # outward


def test_fibonacci_sphere_area_converges():
    errs = []
    for n in (256, 1024, 4096):
        nodes, tri6 = o.fibonacci_sphere_mesh(n, 100.0)
        s = o.CurvedSurface(nodes, tri6)
        errs.append(abs(s.area.sum() / (np.pi * 100.0 ** 2) - 1))
        assert np.allclose(np.linalg.norm(nodes, axis=1), 50.0, rtol=1e-14)
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-5


def _ellipsoid_area(a, b, c, m=400):
    """Surface integral of the ellipsoid by Gauss-Legendre in (t, phi)."""
    t, wt = np.polynomial.legendre.leggauss(m)
    th = 0.5 * np.pi * (t + 1)
    ph = np.pi * (t + 1)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    W = np.outer(wt, wt) * (0.5 * np.pi) * np.pi
    st, ct, sp, cp = np.sin(TH), np.cos(TH), np.sin(PH), np.cos(PH)
    xt = np.stack([a * ct * cp, b * ct * sp, -c * st], -1)
    xp = np.stack([-a * st * sp, b * st * cp, 0 * st], -1)
    return float((np.linalg.norm(np.cross(xt, xp), axis=-1) * W).sum())


def test_triellipsoid_flat_geometry():
    axes = np.array([12.0, 20.0, 33.0])
    nodes, tri6 = o.triellipsoid_mesh(900, axes)
    v0, v1, v2 = nodes[tri6[:, 0]], nodes[tri6[:, 1]], nodes[tri6[:, 2]]
    s = o.CurvedSurface(nodes, tri6)
    cr = np.cross(v0 - v1, v2 - v1)  # particle.py _norm_flat
    # straight-edge quadratic element == flat triangle (particle.py _norm_flat / _quad_flat)
    assert np.allclose(s.pos, (v0 + v1 + v2) / 3, rtol=0, atol=1e-13)
    assert np.allclose(s.area, 0.5 * np.linalg.norm(cr, axis=1), rtol=1e-13)
    assert np.allclose(s.nvec, cr / np.linalg.norm(cr, axis=1, keepdims=True), atol=1e-13)
    # polar quadrature points and weights of particle.py _quadpol_flat
    pts, wts = s.quadpol()
    xq, yq, wq = o.polar_rule_tri()
    flat = (xq[None, :, None] * v0[:, None] + yq[None, :, None] * v1[:, None]
            + (1 - xq - yq)[None, :, None] * v2[:, None])
    assert np.allclose(np.asarray(pts).reshape(flat.shape), flat, atol=1e-13)
    assert np.allclose(np.asarray(wts).reshape(len(v0), -1), wq[None, :] * s.area[:, None], rtol=1e-12)
    # enclosed volume (divergence theorem) and area vs the smooth ellipsoid
    # (inscribed flat polyhedron: errors O(h^2), shrinking ~4x per 4x faces)
    errs = []
    for n in (225, 900):
        nd, t6 = o.triellipsoid_mesh(n, axes)
        sf = o.CurvedSurface(nd, t6)
        vol = np.sum(np.sum(sf.pos * sf.nvec, axis=1) * sf.area) / 3
        errs.append((abs(vol / (4 / 3 * np.pi * axes.prod()) - 1), abs(sf.area.sum() / _ellipsoid_area(*axes) - 1)))
    assert errs[1][0] < 1e-2 and errs[1][1] < 1e-2
    assert errs[0][0] > 3 * errs[1][0] and errs[0][1] > 3 * errs[1][1]


def test_product_meshes_equal_oracle():
    g = pytest.importorskip("mnpbem_amd.geometry")
    for n in (144, 997):
        a = g.fibonacci_sphere(n)
        assert np.array_equal(a, o.golden_spiral(n))
        assert np.array_equal(g.sphere_triangulate(a), o.spherical_delaunay(a))
    p = g.trisphere_fibonacci(1000, 100.0)
    nodes, tri6 = o.fibonacci_sphere_mesh(1000, 100.0)
    assert np.array_equal(p.verts2[p.faces6], nodes[tri6])
    s = o.CurvedSurface(nodes, tri6)
    assert np.allclose(p.pos, s.pos, rtol=0, atol=1e-13) and np.allclose(p.area, s.area, rtol=1e-13)
    q = g.triellipsoid(400, [10.0, 25.0, 40.0])
    nodes, tri6 = o.triellipsoid_mesh(400, [10.0, 25.0, 40.0])
    assert q.interp == "flat"
    assert np.array_equal(q.verts2[q.faces6], nodes[tri6])
    s = o.CurvedSurface(nodes, tri6)
    assert np.allclose(q.pos, s.pos, rtol=0, atol=1e-13) and np.allclose(q.nvec, s.nvec, atol=1e-13)
    # every shipped sphere size loads, nodes bit-identical to the oracle's
    for n in (144, 169, 225, 289, 1024, 1225, 1444):
        t = g.trisphere(n, 10.0)
        nodes, tri6 = o.trisphere_mesh(n, 10.0)
        assert t.n == 2 * n - 4 and np.array_equal(t.verts2[t.faces6], nodes[tri6])


ROD_CASES = {"d20_h60_n15_8_20": (20.0, 60.0, [15, 8, 20]), "d10_h40_n21_11_31": (10.0, 40.0, [21, 11, 31])}


def _golden_rods():
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "trirod_tri.npz"))


@pytest.mark.parametrize("key", sorted(ROD_CASES))
def test_trirod_matches_reference_golden(key):
    """trirod(..., triangles=True) against golden vectors written by the
    reference's own mesh code (tests/golden/make_trirod_golden.py): face order
    and corner coordinates bit-identical; centroids, normals, areas (the
    reference's flat _norm) to round-off."""
    g = pytest.importorskip("mnpbem_amd.geometry")
    z = _golden_rods()
    d, h, n = ROD_CASES[key]
    p = g.trirod(d, h, n, triangles=True)
    assert np.array_equal(p.verts2[p.faces6[:, :3]], z[key + "_corners"])
    assert np.abs(p.pos - z[key + "_pos"]).max() < 1e-13 * d
    assert np.abs(p.nvec - z[key + "_nvec"]).max() < 1e-13
    assert np.abs(p.area / z[key + "_area"] - 1).max() < 1e-13
    # the oracle builds the same surface from the golden corners
    nodes, tri6 = o.flat_mesh_from_corners(z[key + "_corners"])
    s = o.CurvedSurface(nodes, tri6)
    assert np.allclose(s.pos, z[key + "_pos"], rtol=0, atol=1e-13 * d)
    assert np.allclose(s.area, z[key + "_area"], rtol=1e-13)


def test_trirod_closed_outward_and_converging():
    """Closed 2-manifold with outward normals; area and enclosed volume
    converge to the capsule's pi d (h - d) + pi d^2 and
    pi (d/2)^2 (h - d) + 4/3 pi (d/2)^3."""
    g = pytest.importorskip("mnpbem_amd.geometry")
    d, h = 20.0, 60.0
    errs = []
    for n in ([11, 6, 10], [21, 11, 20], [41, 21, 40]):
        p = g.trirod(d, h, n, triangles=True)
        f = p.faces6[:, :3]
        directed = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        assert len({tuple(e) for e in directed}) == len(directed)
        _, cnt = np.unique(np.sort(directed, axis=1), axis=0, return_counts=True)
        assert np.all(cnt == 2)
        c = p.pos
        zc = np.clip(c[:, 2], -(h - d) / 2, (h - d) / 2)
        out = c - np.stack([0 * zc, 0 * zc, zc], 1)   # from the rod axis segment
        assert np.all(np.sum(out * p.nvec, axis=1) > 0)
        vol = np.sum(np.sum(c * p.nvec, axis=1) * p.area) / 3
        a_ref = np.pi * d * (h - d) + np.pi * d ** 2
        v_ref = np.pi * (d / 2) ** 2 * (h - d) + 4 / 3 * np.pi * (d / 2) ** 3
        errs.append((abs(p.area.sum() / a_ref - 1), abs(vol / v_ref - 1)))
    assert errs[0][0] > 3 * errs[1][0] > 9 * errs[2][0] and errs[2][0] < 5e-3
    assert errs[0][1] > 3 * errs[1][1] > 9 * errs[2][1] and errs[2][1] < 5e-3


def test_trirod_quadrilaterals_match_reference():
    """trirod's default quadrilateral rod (triangles=False; the reference's
    rod validation case trirod(10, 40, [15, 15, 15]) with interp='curv'):
    face order, nodes of every face and the corners after clean() bit for bit
    equal to the reference's (tests/golden/make_trirod_quad_golden.py),
    centroids, normals, tangents, areas and element radii to rounding."""
    import os
    import mnpbem_amd as m
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "trirod_quad.npz"))
    rod = m.trirod(10.0, 40.0, [15, 15, 15], triangles=False)
    gn = d["nodes"]
    assert rod.n == len(gn) == 588
    assert np.array_equal(rod.nvert == 4, ~np.isnan(gn[:, 3, 0]))
    ok = ~np.isnan(gn)
    assert np.array_equal(rod.nodes9[ok], gn[ok])
    oc = ~np.isnan(d["corners"])
    assert np.array_equal(rod.corners[oc], d["corners"][oc])
    for key, val in (("pos", rod.pos), ("nvec", rod.nvec), ("tvec1", rod.tvec1), ("tvec2", rod.tvec2),
                     ("area", rod.area), ("brad", rod.bradius())):
        assert np.abs(val - d[key]).max() < 1e-13, key
    # a closed surface: the curved area approaches the rod's 2 pi r (h - d) + pi d^2
    assert abs(rod.area.sum() / (2 * np.pi * 5 * 30 + np.pi * 100) - 1) < 1e-4


def test_translated_copy_detection():
    """The iterative solver's block preconditioner shares one factorisation
    between particles that are shifted copies of each other: shifted spheres
    and rods qualify, a scaled or rotated copy or other materials do not."""
    import mnpbem_amd as m
    from mnpbem_amd.bem_iter import _translated_copy
    p1 = m.trisphere_fibonacci(500, 100.0)
    assert _translated_copy(p1, [2, 1], p1.shift([102.0, 0.0, 0.0]), [2, 1])
    assert not _translated_copy(p1, [2, 1], p1.shift([102.0, 0.0, 0.0]), [3, 1])
    assert not _translated_copy(p1, [2, 1], m.trisphere_fibonacci(500, 90.0), [2, 1])
    rot = m.Particle(p1.verts2[:, [1, 0, 2]], p1.faces6)
    assert not _translated_copy(p1, [2, 1], rot, [2, 1])
    rod = m.trirod(10.0, 40.0, n=[15, 15, 15])
    assert _translated_copy(rod, [2, 1], rod.shift([0.0, 50.0, 0.0]), [2, 1])
    assert not _translated_copy(rod, [2, 1], p1, [2, 1])
