"""Write golden vectors for the reference's default trirod (quadrilateral
faces, triangles=False) as the rod validation case builds it:
trirod(10, 40, n=[15, 15, 15]) in ComParticle(..., interp='curv')
(validation/03_bemret/rod/run_python.py; mnpbem/geometry/mesh_generators.py:
592-640, particle.py:551-613, 1234-1275, 1276-1429).

Run once in a container that holds the reference tree (it imports the
reference's numpy code; nothing of the reference is copied):

    python tests/golden/make_trirod_quad_golden.py /root/reference

Output: tests/golden/trirod_quad.npz with the nodes of every face in face order
(F, 9, 3): corners 0-3, edge midpoints 4-7, centre 8 for quadrilaterals;
corners 0-2 and midpoints 4-6 for triangles, NaN elsewhere), the corner
vertices after the reference's clean() (F, 4, 3), and the reference's
centroids, normals, tangents, areas and element radii of the curved faces.
"""

import os
import sys

import numpy as np


def main(ref_root):
    sys.path.insert(0, ref_root)
    from mnpbem.geometry import ComParticle, trirod
    from mnpbem.materials import EpsConst, EpsTable
    rod = trirod(10.0, 40.0, n=[15, 15, 15])
    cp = ComParticle([EpsConst(1.0), EpsTable("gold.dat")], [rod], [[2, 1]], 1, interp="curv")
    q = cp.p[0]
    f2 = q.faces2
    nodes = np.full((len(f2), 9, 3), np.nan)
    ok = ~np.isnan(f2)
    nodes[ok] = q.verts2[f2[ok].astype(int)]
    f = q.faces
    corners = np.full((len(f), 4, 3), np.nan)
    okc = ~np.isnan(f)
    corners[okc] = q.verts[f[okc].astype(int)]
    out = {"nodes": nodes, "corners": corners, "pos": q.pos, "nvec": q.nvec, "tvec1": q.tvec1, "tvec2": q.tvec2,
           "area": q.area, "brad": q.bradius()}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trirod_quad.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, "faces", len(f2), "quads", int(np.sum(ok[:, 3])))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
