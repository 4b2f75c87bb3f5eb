"""Write golden vectors for trirod(..., triangles=True): the flat-triangle
rod of mnpbem/geometry/mesh_generators.py:592-640 (caps from
trispheresegment, cylinder from fvgrid, merged by Particle.clean).

Run once in a container that holds the reference tree (it imports the
reference's numpy code; nothing of the reference is copied):

    python tests/golden/make_trirod_golden.py /root/reference

Output: tests/golden/trirod_tri.npz with, per case, the corner coordinates of
every face in face order (F, 3, 3) after the reference's clean(), and its
centroids, normals and areas.
"""

import os
import sys

import numpy as np

CASES = {"d20_h60_n15_8_20": (20.0, 60.0, [15, 8, 20]), "d10_h40_n21_11_31": (10.0, 40.0, [21, 11, 31])}


def main(ref_root):
    sys.path.insert(0, ref_root)
    from mnpbem.geometry.mesh_generators import trirod
    out = {}
    for key, (d, h, n) in CASES.items():
        p = trirod(d, h, n=n, triangles=True)
        f = p.faces[:, :3].astype(int)
        assert np.all(np.isnan(p.faces[:, 3]))
        out[key + "_corners"] = p.verts[f]
        out[key + "_pos"] = p.pos
        out[key + "_nvec"] = p.nvec
        out[key + "_area"] = p.area
        print(key, "faces", len(f))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trirod_tri.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
