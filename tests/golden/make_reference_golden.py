"""Write golden vectors from the reference's own code for the parts of the
path that no committed fixture pins (Drude-gold spectra, the refinement pair
list, the Green matrices with refined and closed-surface-corrected entries,
the golden-spiral meshes, the H-matrix block partition).

Run once in a container that holds the reference tree (it imports the
reference's numpy code; nothing of the reference is copied):

    python tests/golden/make_reference_golden.py /root/reference

Output: tests/golden/reference_n284.npz (about 0.25 MB).
  ext, sca          (41, 2)  BEMRet on trisphere(144, 20), EpsConst(1) outside,
                             EpsDrude(10, 9.065, 0.0708) inside, pol x / y, dir +z,
                             lambda = linspace(400, 800, 41)  (config 1, both pols)
  ref_row, ref_col, ref_kind  refinematrix pairs (kind 2 self, 1 near) at N = 284
  green_rows        sampled row indices; G1, H1, G2, H2 rows at 520 nm
                    (CompGreenRet.eval, inside 0 / outside 1)
  fib_nodes         (F, 6, 3) nodes of _trisphere_fibonacci(300, 40) per face
  ell_corners       (F, 3, 3) corners of triellipsoid(400, [10, 25, 40]) per face
  pos               (284, 3) face centroids of the curved trisphere(144, 20)
  tree_*            ClusterTree(pos, cleaf=16) of those centroids and the
                    admissibility list of HMatrix (row1, col1, row2, col2)
"""

import os
import sys

import numpy as np

LAMS = np.linspace(400.0, 800.0, 41)
GREEN_LAM = 520.0
GREEN_ROWS = np.array([0, 1, 7, 33, 64, 100, 141, 142, 200, 255, 283])


def main(ref_root):
    sys.path.insert(0, ref_root)
    from mnpbem.materials import EpsConst, EpsDrude
    from mnpbem.geometry import ComParticle
    from mnpbem.geometry.mesh_generators import trisphere, _trisphere_fibonacci, triellipsoid
    from mnpbem.greenfun import CompGreenRet
    from mnpbem.greenfun.refine_utils import refinematrix
    from mnpbem.greenfun.clustertree import ClusterTree
    from mnpbem.greenfun.hmatrix import HMatrix
    from mnpbem.bem import BEMRet
    from mnpbem.simulation import PlaneWaveRet

    out = {}
    eps = [EpsConst(1.0), EpsDrude(10.0, 9.065, 0.0708)]
    p = ComParticle(eps, [trisphere(144, 20.0)], [[2, 1]], 1, interp="curv")
    print("N =", p.n, flush=True)

    # refinement pairs (refine_utils.py:29-158)
    ir = refinematrix(p, p)
    ir = ir.tocoo() if hasattr(ir, "tocoo") else ir
    order = np.lexsort((ir.col, ir.row))
    out["ref_row"] = ir.row[order].astype(np.int32)
    out["ref_col"] = ir.col[order].astype(np.int32)
    out["ref_kind"] = np.asarray(ir.data)[order].astype(np.int8)

    # Green matrices, sampled rows
    g = CompGreenRet(p, p)
    out["green_rows"] = GREEN_ROWS
    for name, (i, key) in {"G1": (0, "G"), "H1": (0, "H1"), "G2": (1, "G"), "H2": (1, "H2")}.items():
        M = np.asarray(g.eval(i, i, key, GREEN_LAM))
        out[name] = M[GREEN_ROWS]
    print("green done", flush=True)

    # spectrum (config 1, both polarisations)
    exc = PlaneWaveRet(np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]), np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]))
    bem = BEMRet(p)
    ext = np.zeros((len(LAMS), 2))
    sca = np.zeros((len(LAMS), 2))
    for k, lam in enumerate(LAMS):
        sig, bem = bem.solve(exc.potential(p, lam))
        ext[k] = np.real(np.ravel(exc.extinction(sig)))
        s = exc.scattering(sig)
        sca[k] = np.real(np.ravel(s[0] if isinstance(s, tuple) else s))
    out["lams"], out["ext"], out["sca"] = LAMS, ext, sca
    print("spectrum done", flush=True)

    # golden-spiral meshes
    fp = _trisphere_fibonacci(300, 40.0)
    out["fib_nodes"] = fp.verts2[fp.faces2[:, [0, 1, 2, 4, 5, 6]].astype(int)]
    ep = triellipsoid(400, [10.0, 25.0, 40.0])
    out["ell_corners"] = ep.verts[ep.faces[:, :3].astype(int)]

    # cluster tree and block partition of the N = 284 centroids
    pos = p.pos
    out["pos"] = pos
    t = ClusterTree(pos, cleaf=16)
    out["tree_son"], out["tree_cind"], out["tree_ind"] = t.son, t.cind, t.ind
    out["tree_mid"], out["tree_rad"] = t.mid, t.rad
    h = HMatrix(tree=t, htol=1e-6, kmax=50)
    out["hm_row1"], out["hm_col1"] = np.asarray(h.row1), np.asarray(h.col1)
    out["hm_row2"], out["hm_col2"] = np.asarray(h.row2), np.asarray(h.col2)

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_n284.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
