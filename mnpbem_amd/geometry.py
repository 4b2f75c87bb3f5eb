"""Particle surfaces of quadratic (6-node) triangles, the sphere mesh
generator and the composite particle (host-side set-up).

Mirrors mnpbem/geometry: ``trisphere`` (mesh_generators.py:13-130), the
golden-spiral meshes ``trisphere_fibonacci`` and ``triellipsoid``
(mesh_generators.py:310-473), ``Particle`` with curved or flat interpolation
(particle.py:310-613; a flat triangle is carried as the quadratic element
whose edge nodes are the straight-edge midpoints) and ``ComParticle``
(comparticle.py:55-127).  Only what the retarded BEM path
needs is kept: centroids, normals, tangents, areas, element radii and the
quadrature nodes that the GPU refinement kernel integrates over.
"""

import os

import numpy as np

from .quadrature import rule18, polar_rule, tri6_tables

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# vertex counts of the saved sphere point sets (mesh_generators.py:58-60)
SPHERE_SIZES = np.array([32, 60, 144, 169, 225, 256, 289, 324, 361, 400, 441, 484,
                         529, 576, 625, 676, 729, 784, 841, 900, 961, 1024, 1225, 1444])


def _read_sphere(n):
    """int32 nv, nf; float64 x[nv], y[nv], z[nv]; int32 faces[nf, 3] (1-based)."""
    path = os.path.join(_DATA, "sphere%d_tri.bin" % n)
    if not os.path.exists(path):
        raise FileNotFoundError("sphere mesh with %d vertices is not shipped (%s)" % (n, path))
    raw = open(path, "rb").read()
    nv, nf = np.frombuffer(raw, dtype=np.int32, count=2)
    off = 8
    xyz = np.frombuffer(raw, dtype=np.float64, count=3 * nv, offset=off).reshape(3, nv).T.copy()
    off += 24 * nv
    faces = np.frombuffer(raw, dtype=np.int32, count=3 * nf, offset=off).reshape(nf, 3).astype(np.int64) - 1
    return xyz, faces


class Particle:
    """Surface of curved triangles given by 6 nodes per face: three corners
    and the three edge midpoints (v0, v1, v2, m01, m12, m20).

    Attributes (numpy, float64): ``pos`` centroids (n, 3), ``nvec`` unit
    normals, ``tvec1``/``tvec2`` tangents, ``area``, ``nodes`` (n, 6, 3).
    """

    def __init__(self, verts2, faces6, interp="curv"):
        if interp not in ("curv", "flat"):
            raise ValueError("Particle: interp must be 'curv' or 'flat'")
        verts2 = np.asarray(verts2, dtype=float)
        faces6 = np.asarray(faces6, dtype=np.int64)
        if verts2.ndim != 2 or verts2.shape[1] != 3:
            raise ValueError("Particle: 'verts' must have shape (N, 3); got %s." % (verts2.shape,))
        self.verts = self.faces = self.mid_edges = None
        if faces6.ndim == 2 and faces6.shape[1] == 3:
            # plain triangles: edge midpoints on the straight edges, so the
            # quadratic element is the flat triangle itself (its shape
            # functions reproduce the linear map exactly)
            self.verts, self.faces = verts2, faces6
            verts2, faces6, self.mid_edges = add_midpoints(verts2, faces6, with_edges=True)
        if faces6.ndim != 2 or faces6.shape[1] != 6:
            raise ValueError("Particle: faces must list 3 corners or 6 nodes per triangle")
        self.verts2 = verts2
        self.faces6 = faces6
        self.interp = interp
        self._geometry()

    @property
    def n(self):
        return len(self.faces6)

    nfaces = n

    @property
    def nvec(self):
        return self._nvec

    @property
    def tvec1(self):
        return self._t1

    @property
    def tvec2(self):
        return self._t2

    def _geometry(self):
        X = self.verts2[self.faces6]                                  # (n, 6, 3)
        self.nodes = np.ascontiguousarray(X)
        # area = integral of the surface Jacobian with the 28-point rule
        S, Dx, Dy, w = tri6_tables(*rule18())
        tx = np.einsum("qc,ncd->nqd", Dx, X)
        ty = np.einsum("qc,ncd->nqd", Dy, X)
        jac = np.linalg.norm(np.cross(tx, ty), axis=2)
        self.area = (0.5 * w[None, :] * jac).sum(axis=1)
        # centroid and tangents from the shape functions at (1/3, 1/3)
        S0, Dx0, Dy0, _ = tri6_tables(np.array([1 / 3]), np.array([1 / 3]), np.array([1.0]))
        self.pos = X.transpose(0, 2, 1) @ S0[0]
        a = X.transpose(0, 2, 1) @ Dy0[0]
        b = X.transpose(0, 2, 1) @ Dx0[0]
        nv = np.cross(a, b)
        nv /= np.maximum(np.linalg.norm(nv, axis=1, keepdims=True), 1e-14)
        a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-14)
        if self.interp == "flat":
            # particle.py _norm_flat: tvec1 along the edge from corner 1 to corner 0
            a = X[:, 0, :] - X[:, 1, :]
            a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-14)
        self._nvec, self._t1 = nv, a
        self._t2 = np.cross(nv, a)
        corners = X[:, :3, :]
        self._brad = np.sqrt(((corners - self.pos[:, None, :]) ** 2).sum(axis=2)).max(axis=1)
        # the 9-node layout of mixed triangle / quadrilateral surfaces
        # (corners 0-2, midpoints 4-6 for a triangle)
        self.nodes9 = np.zeros((self.n, 9, 3))
        self.nodes9[:, [0, 1, 2, 4, 5, 6]] = X
        self.nvert = np.full(self.n, 3, dtype=np.int32)

    def shift(self, vec):
        """Copy of the particle translated by ``vec`` (particle.py shift)."""
        return Particle(self.verts2 + np.asarray(vec, dtype=float).reshape(1, 3), self.faces6, interp=self.interp)

    def bradius(self):
        """Radius of the smallest centroid-centred sphere through the corners
        (particle.py:2848-2880)."""
        return self._brad


class MixedParticle(Particle):
    """Curved surface of 6-node triangles and 9-node quadrilaterals
    (particle.py:551-613, 1234-1429): ``nodes9`` (n, 9, 3) holds the corners
    (0-3), edge midpoints (4-7) and centre (8) of a quadrilateral, or the
    corners (0-2) and midpoints (4-6) of a triangle; ``nvert`` (n,) is 3 or
    4; ``corners`` (n, 4, 3) the corner vertices of the cleaned mesh (NaN in
    slot 3 of a triangle), from which the element radii are taken.

    A quadrilateral integrates as its two 6-node triangles (v0, v1, v2, m01,
    m12, c) and (v2, v3, v0, m23, m30, c) with the 28-point rule
    (particle.py:695-720, 1520-1575); its centroid is the centre node and its
    tangents come from the first triangle (particle.py:582-596); its self
    term uses the 9-node shape functions with the quadrilateral polar rule."""

    def __init__(self, nodes9, nvert, corners, interp="curv"):
        self.nodes9 = np.ascontiguousarray(np.asarray(nodes9, dtype=float))
        self.nvert = np.ascontiguousarray(np.asarray(nvert, dtype=np.int32))
        self.corners = np.asarray(corners, dtype=float)
        self.interp = interp
        self.verts = self.faces = self.mid_edges = None
        self.faces6 = np.zeros((len(self.nvert), 6), dtype=np.int64)   # face count only
        self._geometry()

    def _geometry(self):
        N9, q = self.nodes9, self.nvert == 4
        sub1 = N9[:, [0, 1, 2, 4, 5, 6]].copy()
        sub1[q] = N9[q][:, [0, 1, 2, 4, 5, 8]]
        sub2 = N9[q][:, [2, 3, 0, 6, 7, 8]]
        self.nodes = np.ascontiguousarray(sub1)
        S, Dx, Dy, w = tri6_tables(*rule18())
        jac = lambda X: np.linalg.norm(np.cross(np.einsum("qc,ncd->nqd", Dx, X),
                                                np.einsum("qc,ncd->nqd", Dy, X)), axis=2)
        self.area = (0.5 * w[None, :] * jac(sub1)).sum(axis=1)
        self.area[q] += (0.5 * w[None, :] * jac(sub2)).sum(axis=1)
        S0, Dx0, Dy0, _ = tri6_tables(np.array([1 / 3]), np.array([1 / 3]), np.array([1.0]))
        pos = sub1.transpose(0, 2, 1) @ S0[0]
        a = sub1.transpose(0, 2, 1) @ Dy0[0]
        b = sub1.transpose(0, 2, 1) @ Dx0[0]
        pos[q] = N9[q, 8]
        a[q] = sub1[q].transpose(0, 2, 1) @ np.array([0.0, -1.0, -1.0, 2.0, 2.0, -2.0])
        b[q] = sub1[q].transpose(0, 2, 1) @ np.array([1.0, 0.0, -1.0, 0.0, 0.0, 0.0])
        nv = np.cross(a, b)
        nv /= np.maximum(np.linalg.norm(nv, axis=1, keepdims=True), 1e-14)
        a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-14)
        self.pos, self._nvec, self._t1 = pos, nv, a
        self._t2 = np.cross(nv, a)
        d = np.sqrt(((self.corners - pos[:, None, :]) ** 2).sum(axis=2))
        self._brad = np.nanmax(d, axis=1)

    @property
    def n(self):
        return len(self.nvert)

    nfaces = n

    def shift(self, vec):
        v = np.asarray(vec, dtype=float).reshape(1, 1, 3)
        return MixedParticle(self.nodes9 + v, self.nvert, self.corners + v, interp=self.interp)


def _quad_grid(m, n):
    """Cells of an (m, n) vertex grid (vertices in column-major order) as
    quadrilaterals [v00, v01, v11, v10], columns outer, rows inner."""
    jj, ii = np.meshgrid(np.arange(n - 1), np.arange(m - 1), indexing="ij")
    v00 = (jj * m + ii).ravel()
    return np.stack([v00, v00 + m, v00 + m + 1, v00 + 1], axis=1)


def _nodes9(verts, faces4):
    """9-node layout of triangles (faces4[:, 3] < 0) and quadrilaterals:
    corners, edge midpoints 0.5 (a + b), centre 0.25 (v0 + v1 + v2 + v3)
    (mesh_generators.py:133-196)."""
    tri = faces4[:, 3] < 0
    N = np.zeros((len(faces4), 9, 3))
    for k in range(4):
        N[:, k] = verts[np.where(faces4[:, k] < 0, 0, faces4[:, k])]
    N[tri, 3] = 0.0
    for e, (a, b) in enumerate(((0, 1), (1, 2), (2, 3), (3, 0))):
        N[:, 4 + e] = 0.5 * (N[:, a] + N[:, b])
    N[tri, 6] = 0.5 * (N[tri, 2] + N[tri, 0])
    N[tri, 7] = 0.0
    N[:, 8] = 0.25 * (N[:, 0] + N[:, 1] + N[:, 2] + N[:, 3])
    N[tri, 8] = 0.0
    return N


def _clean_quads(verts, faces):
    """Particle.clean (particle.py:2331-2388) for quadrilateral grids: merge
    vertices equal after rounding to 8 decimals (they take the rounded,
    sorted coordinates); a quadrilateral left with three distinct corners
    becomes the triangle of those corners in their order (slot 3 = -1)."""
    vr = _round_digits(verts, 8)
    uniq, inv = np.unique(vr, axis=0, return_inverse=True)
    f = inv.reshape(-1)[faces]
    out = np.full((len(f), 4), -1, dtype=np.int64)
    for i, row in enumerate(f):
        _, first = np.unique(row, return_index=True)
        keep = row[np.sort(first)]
        out[i, :len(keep)] = keep
    return uniq, out


def _trirod_quads(d, h, phi, theta, nz):
    """The reference's default rod: quadrilateral faces with triangles at the
    poles (mesh_generators.py:549-640).  Cap = sphere segment over (theta,
    phi): grid cells merged (the pole row collapses to triangles), flat
    midpoints and centres from the merged corners, every node projected onto
    the sphere, shifted up by (h - d)/2; lower cap = its mirror image in z
    with reversed faces; cylinder = (phi, z) grid cells in reversed corner
    order with midpoints and centres taken in the (phi, z) parameters and
    mapped onto the cylinder.  Corners of the three parts are merged once
    more (rounded to 8 decimals) for the element radii."""
    pg, tg = np.meshgrid(phi, theta)
    cx, cy, cz = d / 2.0 * np.sin(tg) * np.cos(pg), d / 2.0 * np.sin(tg) * np.sin(pg), d / 2.0 * np.cos(tg)
    cv = np.column_stack([cx.ravel(order="F"), cy.ravel(order="F"), cz.ravel(order="F")])
    cv, cf = _clean_quads(cv, _quad_grid(*cx.shape))
    N1 = _nodes9(cv, cf)
    tri = cf[:, 3] < 0
    nrm = np.maximum(np.sqrt(np.sum(N1 ** 2, axis=2, keepdims=True)), 1e-30)
    N1 = 0.5 * d * (N1 / nrm)
    N1[tri, 3] = N1[tri, 7] = N1[tri, 8] = 0.0
    N1 = N1 + np.array([0.0, 0.0, 0.5 * (h - d)])
    N2 = N1 * np.array([1.0, 1.0, -1.0])
    N2q = N2[~tri][:, [3, 2, 1, 0, 6, 5, 4, 7, 8]]
    N2t = N2[tri][:, [2, 1, 0, 3, 5, 4, 6, 7, 8]]
    N2 = N2.copy()
    N2[~tri], N2[tri] = N2q, N2t
    zv = 0.5 * _mlinspace(-1, 1, nz) * (h - d)
    gx, gy = np.meshgrid(phi, zv)
    gv = np.column_stack([gx.ravel(order="F"), gy.ravel(order="F"), np.zeros(gx.size)])
    gf = _quad_grid(*gx.shape)[:, ::-1]
    Ng = _nodes9(gv, gf)
    N3 = np.stack([0.5 * d * np.cos(Ng[:, :, 0]), 0.5 * d * np.sin(Ng[:, :, 0]), Ng[:, :, 1]], axis=2)
    nodes = np.concatenate([N1, N2, N3])
    nvert = np.concatenate([np.where(tri, 3, 4), np.where(tri, 3, 4), np.full(len(gf), 4)]).astype(np.int32)
    # corners of the merged mesh (element radii)
    cn = nodes[:, :4].copy()
    cn[nvert == 3, 3] = np.nan
    flat = cn.reshape(-1, 3)
    ok = ~np.isnan(flat[:, 0])
    flat[ok] = _round_digits(flat[ok], 8)
    return MixedParticle(nodes, nvert, flat.reshape(-1, 4, 3))


def add_midpoints(verts, faces, with_edges=False):
    """Append one node per edge at the edge midpoint: returns (nodes,
    faces6) with faces6 = (v0, v1, v2, m01, m12, m20)
    (mesh_generators.py:133-196); with ``with_edges`` also the (nedges, 2)
    corner pair of each appended node."""
    verts = np.asarray(verts, dtype=float)
    faces = np.asarray(faces, dtype=np.int64)
    edges = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]), axis=1)
    uniq, inv = np.unique(edges, axis=0, return_inverse=True)
    mid = 0.5 * (verts[uniq[:, 0]] + verts[uniq[:, 1]])
    nodes = np.vstack([verts, mid])
    faces6 = np.hstack([faces, len(verts) + inv.reshape(3, -1).T])
    return (nodes, faces6, uniq) if with_edges else (nodes, faces6)


def fibonacci_sphere(n):
    """n points of the golden-angle spiral on the unit sphere
    (mesh_generators.py:341-370; the same arithmetic, so the points and the
    triangulations built on them are identical)."""
    k = np.arange(0, n, dtype=float) + 0.5
    polar = np.arccos(1 - 2 * k / n)
    azim = np.pi * (1 + 5 ** 0.5) * k
    pts = np.column_stack([np.sin(polar) * np.cos(azim), np.sin(polar) * np.sin(azim), np.cos(polar)])
    return pts / np.linalg.norm(pts, axis=1, keepdims=True)


def sphere_triangulate(verts):
    """Triangulate points on the unit sphere (mesh_generators.py:373-443,
    MATLAB sphtriangulate.m): rotate the first point to the south pole,
    project the others stereographically onto the plane, take their
    Delaunay triangulation (oriented outward) and close the surface with a
    fan from the first point to the planar convex hull."""
    from scipy.spatial import ConvexHull, Delaunay
    verts = np.asarray(verts, dtype=float)
    n = len(verts)
    c = verts[0]
    e3 = -c
    e2 = np.array([0, -e3[2], e3[1]]) if c[2] != 0 else np.array([-e3[1], e3[0], 0])
    e2 = e2 / np.linalg.norm(e2)
    e1 = np.cross(e3, e2)
    rot = np.array([e1, e2, e3])
    vr = verts[1:] @ rot.T
    t = -1.0 / (vr[:, 2] + 1)
    xy = np.column_stack([vr[:, 0] * t, vr[:, 1] * t])
    tri = Delaunay(xy).simplices
    p3 = np.column_stack([xy, np.zeros(n - 1)])
    w = np.cross(p3[tri[:, 0]] - p3[tri[:, 1]], p3[tri[:, 2]] - p3[tri[:, 1]])
    flip = w[:, 2] > 0
    tri[flip, :] = tri[flip, :][:, [0, 2, 1]]
    hv = ConvexHull(xy).vertices
    nh = len(hv)
    fan = np.array([[hv[(i + 1) % nh] + 1, hv[i] + 1, 0] for i in range(nh)], dtype=np.int64)
    return np.vstack([tri + 1, fan]).astype(np.int64)


def trisphere_fibonacci(n, diameter=1.0, **kwargs):
    """Sphere of curved triangles on n golden-spiral points: Delaunay on the
    sphere, midpoints added and all nodes projected onto the sphere
    (mesh_generators.py:310-338)."""
    verts = fibonacci_sphere(int(n))
    faces = sphere_triangulate(verts)
    nodes, faces6 = add_midpoints(verts * (diameter / 2.0), faces)
    nodes = (diameter / 2.0) * (nodes / np.linalg.norm(nodes, axis=1, keepdims=True))
    kwargs.setdefault("interp", "curv")
    return Particle(nodes, faces6, **kwargs)


def triellipsoid(n, axes):
    """Ellipsoid of flat triangles with semi-axes ``axes`` on n golden-spiral
    points (mesh_generators.py:446-473)."""
    axes = np.asarray(axes, dtype=float).reshape(3)
    if np.any(axes <= 0):
        raise ValueError("triellipsoid: semi-axes must be positive")
    verts = fibonacci_sphere(int(n))
    faces = sphere_triangulate(verts)
    return Particle(verts * axes, faces, interp="flat")


def _mlinspace(a, b, n):
    """a + i (b - a) / (n - 1) with the last point exactly b (the MATLAB
    linspace of mnpbem/utils/matlab_compat.py mlinspace)."""
    a, b = float(a), float(b)
    if n == 1:
        return np.array([b])
    y = a + np.arange(n, dtype=np.float64) * (b - a) / (n - 1)
    y[-1] = b
    return y


def _round_digits(x, k):
    """Round half away from zero to k decimals: round(x 10^k) / 10^k."""
    y = np.asarray(x, dtype=np.float64) * (10.0 ** k)
    return np.trunc(y + np.copysign(0.5, y)) / (10.0 ** k)


def _grid_triangles(x, y, z):
    """Vertices (column-major order of the (m, n) grids) and the two triangles
    of every grid cell: all [v00, v01, v11] first, then all [v00, v11, v10]
    (cell (i, j) with v00 = j m + i, v10 = v00 + 1, v01 = v00 + m)."""
    m, n = x.shape
    verts = np.column_stack([x.ravel(order="F"), y.ravel(order="F"), z.ravel(order="F")])
    jj, ii = np.meshgrid(np.arange(n - 1), np.arange(m - 1), indexing="ij")
    v00 = (jj * m + ii).ravel()
    v10, v01, v11 = v00 + 1, v00 + m, v00 + m + 1
    faces = np.concatenate([np.stack([v00, v01, v11], 1), np.stack([v00, v11, v10], 1)])
    return verts, faces


def _merge_clean(verts, faces):
    """Merge vertices equal after rounding to 8 decimals (the merged vertex
    takes the rounded coordinates, vertices sorted) and drop faces whose area
    falls below 1e-10 of the mean (collapsed faces), keeping the face order
    (Particle.clean, particle.py:2331-2388)."""
    vr = _round_digits(verts, 8)
    uniq, inv = np.unique(vr, axis=0, return_inverse=True)
    faces = inv.reshape(-1)[faces]
    c = uniq[faces]
    area = 0.5 * np.linalg.norm(np.cross(c[:, 0] - c[:, 1], c[:, 2] - c[:, 1]), axis=1)
    return uniq, faces[area > 1e-10 * area.mean()]


def trirod(diameter, height, n=None, triangles=False, **kwargs):
    """Rod (cylinder with hemispherical caps) of flat triangles, the face
    order and vertex arithmetic of mesh_generators.py:592-640:
    n = [nphi, ntheta, nz]; upper cap = sphere segment over phi x theta in
    [0, 2 pi] x [0, pi/2] (grid triangles, merged, vertices projected onto
    the sphere) shifted up by (height - diameter)/2; lower cap = its mirror
    image in z with reversed faces; cylinder = phi x z grid triangles with
    reversed corner order; the three are merged (vertices rounded to 8
    decimals) and collapsed faces dropped.  triangles=False (the reference's
    default) gives the curved quadrilateral rod of ``_trirod_quads``."""
    n = [15, 20, 20] if n is None else list(n)
    if len(n) != 3:
        raise ValueError("trirod: n must be [nphi, ntheta, nz]")
    nphi, ntheta, nz = (int(v) for v in n)
    d, h = float(diameter), float(height)
    if not (d > 0 and h >= d):
        raise ValueError("trirod: need diameter > 0 and height >= diameter")
    phi = _mlinspace(0, 2 * np.pi, nphi)
    theta = _mlinspace(0, 0.5 * np.pi, ntheta)
    if not triangles:
        return _trirod_quads(d, h, phi, theta, nz)
    # cap: sphere segment (trispheresegment, mesh_generators.py:549-589)
    pg, tg = np.meshgrid(phi, theta)
    cx = d / 2.0 * np.sin(tg) * np.cos(pg)
    cy = d / 2.0 * np.sin(tg) * np.sin(pg)
    cz = d / 2.0 * np.cos(tg)
    cv, cf = _grid_triangles(cx, cy, cz)
    cv, cf = _merge_clean(cv, cf)
    norms = np.maximum(np.sqrt(np.sum(cv ** 2, axis=1, keepdims=True)), 1e-30)
    cv = 0.5 * d * (cv / norms)
    cap1 = cv + np.array([0.0, 0.0, 0.5 * (h - d)])
    cap2 = cap1 * np.array([1.0, 1.0, -1.0])
    # cylinder (fvgrid, mesh_generators.py:515-546, corner order reversed)
    zv = 0.5 * _mlinspace(-1, 1, nz) * (h - d)
    gx, gy = np.meshgrid(phi, zv)
    gv, gf = _grid_triangles(gx, gy, np.zeros_like(gx))
    gf = gf[:, ::-1]
    cyl = np.column_stack([0.5 * d * np.cos(gv[:, 0]), 0.5 * d * np.sin(gv[:, 0]), gv[:, 1]])
    verts = np.concatenate([cap1, cap2, cyl])
    o2, o3 = len(cap1), 2 * len(cap1)
    faces = np.concatenate([cf, cf[:, ::-1] + o2, gf + o3])
    verts, faces = _merge_clean(verts, faces)
    return Particle(verts, faces, interp="flat")


def trisphere(n, diameter=1.0, **kwargs):
    """Sphere of curved triangles from the saved point set closest to ``n``
    vertices: scale to the diameter, add one node per edge, project all
    nodes onto the sphere (mesh_generators.py:13-130)."""
    if n is None:
        raise ValueError("trisphere: 'n' must be a positive integer, got None.")
    n = int(n)
    if n <= 0:
        raise ValueError("trisphere: 'n' must be > 0, got %d." % n)
    if not np.isfinite(diameter) or diameter <= 0:
        raise ValueError("trisphere: 'diameter' must be a positive finite float, got %r." % diameter)
    m = int(SPHERE_SIZES[np.argmin(np.abs(SPHERE_SIZES - n))])
    verts, faces = _read_sphere(m)
    # same arithmetic order as mesh_generators.py:104-118, so nodes are bit-identical
    nodes, faces6 = add_midpoints(0.5 * verts * diameter, faces)
    nodes = 0.5 * diameter * (nodes / np.linalg.norm(nodes, axis=1, keepdims=True))
    return Particle(nodes, faces6, **kwargs)


class ComParticle:
    """Composite of particles with materials inside/outside.

    ``ComParticle(eps, [p1, p2, ...], [[in, out], ...], *closed)`` with
    1-based material indices (comparticle.py:55-127).  Every particle is
    treated as its own closed surface with outward normals.
    """

    def __init__(self, eps, particles, inout, *closed, **kwargs):
        if eps is None:
            raise ValueError("ComParticle: 'eps' must be a list of dielectric functions, got None.")
        if particles is None:
            raise ValueError("ComParticle: 'particles' must be a list of Particle, got None.")
        if inout is None:
            raise ValueError("ComParticle: 'inout' must be an inside/outside index table, got None.")
        if not hasattr(eps, "__len__") or len(eps) == 0:
            raise ValueError("ComParticle: 'eps' must contain at least one dielectric function.")
        if isinstance(particles, Particle):
            particles = [particles]
        self.eps = list(eps)
        self.p = list(particles)
        io = np.atleast_2d(np.asarray(inout))
        if io.ndim != 2 or io.shape[1] != 2 or io.shape[0] != len(self.p):
            raise ValueError("ComParticle: 'inout' must have shape (nparticles, 2); got %s." % (io.shape,))
        if int(io.max()) > len(self.eps) or int(io.min()) < 1:
            raise ValueError("ComParticle: 'inout' indices are 1-based and must reference eps.")
        if kwargs.get("interp", "curv") != "curv":
            raise NotImplementedError("only interp='curv' is supported")
        self.inout = io.astype(int)
        self.closed = list(closed) if closed else [None] * len(self.p)
        offs = np.cumsum([0] + [q.n for q in self.p])
        self.offsets = offs
        cat = lambda f: np.ascontiguousarray(np.concatenate([f(q) for q in self.p]))
        self.pos = cat(lambda q: q.pos)
        self.nvec = cat(lambda q: q.nvec)
        self.tvec1 = cat(lambda q: q.tvec1)
        self.tvec2 = cat(lambda q: q.tvec2)
        self.area = cat(lambda q: q.area)
        self.nodes = cat(lambda q: q.nodes)
        self.nodes9 = cat(lambda q: q.nodes9)
        self.nvert = cat(lambda q: q.nvert)
        self.brad = cat(lambda q: q.bradius())
        self.inout_faces = np.concatenate([np.tile(self.inout[i], (q.n, 1)) for i, q in enumerate(self.p)])
        self.surface = np.concatenate([np.full(q.n, i) for i, q in enumerate(self.p)])

    @property
    def n(self):
        return int(self.offsets[-1])

    nfaces = n

    @property
    def np(self):
        return len(self.p)

    def index(self, i):
        """0-based face indices of particle i (1-based, as in MATLAB)."""
        return np.arange(self.offsets[i - 1], self.offsets[i])

    def eps1(self, enei):
        return np.array([self.eps[m - 1](enei)[0] for m in self.inout_faces[:, 0]])

    def eps2(self, enei):
        return np.array([self.eps[m - 1](enei)[0] for m in self.inout_faces[:, 1]])
