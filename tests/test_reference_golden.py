"""CPU pins of the oracle (and of the product's host-side mesh and tree code)
against golden vectors written by the reference's own code
(tests/golden/make_reference_golden.py, reference_n284.npz): the
Drude-gold spectrum of config 1 (both polarisations), the refinement pair
list, sampled rows of the four Green matrices (refined near-field entries and
the closed-surface correction included), the golden-spiral sphere and
ellipsoid meshes, the cluster tree and the H-matrix block partition.
"""

import os

import numpy as np
import pytest

import oracle as o
from oracle.hmatrix_oracle import OracleClusterTree

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DRUDE = (10.0, 9.065, 0.0708)


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLD, "reference_n284.npz"))


@pytest.fixture(scope="module")
def drude284():
    nodes, tri6 = o.trisphere_mesh(144, 20.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(*DRUDE)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    return comp, o.GreenSetup(comp)


def test_centroids_match_reference(ref, drude284):
    comp, _ = drude284
    assert np.abs(comp.parts[0].pos - ref["pos"]).max() < 1e-13 * 10.0


def test_refinement_pairs_match_reference(ref, drude284):
    """refinematrix (refine_utils.py:29-158): identical pair list and kinds."""
    comp, _ = drude284
    s = comp.parts[0]
    ir = o.refinematrix(s.pos, s.pos, s.bradius, s.bradius)
    r, c = np.nonzero(ir)
    assert np.array_equal(r, ref["ref_row"]) and np.array_equal(c, ref["ref_col"])
    assert np.array_equal(ir[r, c].astype(np.int8), ref["ref_kind"])


def test_green_rows_match_reference(ref, drude284):
    """Rows of G1, H1, G2, H2 at 520 nm (CompGreenRet.eval): far entries,
    refined near-field entries and the self terms with the closed-surface
    correction.  1e-10: the self terms' polar nodes sit at 5e-7 of the element
    size from the centroid, where |centroid - node| loses ~9 digits."""
    _, setup = drude284
    rows = ref["green_rows"]
    for name, M in zip(("G1", "H1", "G2", "H2"), o.green_matrices(setup, 520.0)):
        want = ref[name]
        assert np.abs(M[rows] - want).max() < 1e-10 * np.abs(want).max(), name


def test_drude_spectrum_matches_reference(ref, drude284):
    """Config 1 (Drude gold, trisphere(144, 20), x and y polarisation, +z):
    the oracle's extinction and scattering against the reference BEMRet on
    every fourth of the 41 wavelengths."""
    comp, setup = drude284
    sel = np.arange(0, 41, 4)
    lams = ref["lams"][sel]
    ext, sca = o.solve_spectrum(comp, lams, [[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]], with_sca=True,
                                setup=setup)
    assert np.max(np.abs(ext / ref["ext"][sel] - 1)) < 1e-10
    assert np.max(np.abs(sca / ref["sca"][sel] - 1)) < 1e-10


def test_golden_spiral_meshes_match_reference(ref):
    """_trisphere_fibonacci(300, 40) and triellipsoid(400, [10, 25, 40]):
    oracle and product meshes bit-identical to the reference's (face order,
    corner and midpoint nodes)."""
    nodes, tri6 = o.fibonacci_sphere_mesh(300, 40.0)
    assert np.array_equal(nodes[tri6], ref["fib_nodes"])
    nodes, tri6 = o.triellipsoid_mesh(400, [10.0, 25.0, 40.0])
    assert np.array_equal(nodes[tri6[:, :3]], ref["ell_corners"])
    g = pytest.importorskip("mnpbem_amd.geometry")
    p = g.trisphere_fibonacci(300, 40.0)
    assert np.array_equal(p.verts2[p.faces6], ref["fib_nodes"])
    q = g.triellipsoid(400, [10.0, 25.0, 40.0])
    assert np.array_equal(q.verts2[q.faces6[:, :3]], ref["ell_corners"])


def test_cluster_tree_and_partition_match_reference(ref):
    """ClusterTree(pos, cleaf=16) of the reference's centroids: oracle and
    product trees node for node; the product's HMatrix block partition in
    the reference's order."""
    pos = ref["pos"]
    t = OracleClusterTree(pos, cleaf=16)
    for name in ("son", "cind", "ind"):
        assert np.array_equal(getattr(t, name), ref["tree_" + name]), name
    assert np.array_equal(t.mid, ref["tree_mid"]) and np.array_equal(t.rad, ref["tree_rad"])
    hm = pytest.importorskip("mnpbem_amd.hmatrix")
    tp = hm.ClusterTree(pos, cleaf=16)
    for name in ("son", "cind", "ind"):
        assert np.array_equal(getattr(tp, name), ref["tree_" + name]), name
    h = hm.HMatrix(tp, htol=1e-6, kmax=50, device="cpu")
    for name in ("row1", "col1", "row2", "col2"):
        assert np.array_equal(getattr(h, name), ref["hm_" + name]), name
