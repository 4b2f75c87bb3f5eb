"""GPU reproduction of the reference's rod validation case: quadrilateral
faces with curved interpolation (validation/03_bemret/rod/run_python.py:
trirod(10, 40, n=[15, 15, 15]) - 560 nine-node quadrilaterals and 28 six-node
triangles - tabulated gold (gold.dat) in vacuum, plane waves polarised along
x (propagating along z) and along z (propagating along x), 41 wavelengths
400-800 nm).  The mesh itself is pinned bit for bit on the CPU
(tests/test_meshes.py).

Pins: the MATLAB spectra of the case (validation/03_bemret/rod/data/
bemret_matlab.csv, 11 significant digits, all 41 wavelengths) and the values
the reference's present Python code computes at 400, 600 and 800 nm
(tests/golden/rod_quad_spectrum.npz, written by
tests/golden/make_rod_quad_spectrum_golden.py); the two agree to 1e-11.  The
committed Python CSV of the case (bemret_python.csv) was written by an
earlier state of the reference and differs from both by up to 6.7e-6, so it
is kept only as data.  Bars: complex128 2e-10, mixed 1e-9 (extinction and
scattering)."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("precision,tol", [("double", 2e-10), ("mixed", 1e-9)])
def test_quadrilateral_rod_spectra_match_reference(cuda, precision, tol):
    import mnpbem_amd as m
    ml = np.loadtxt(os.path.join(GOLD, "rod_d10_h40_n15_goldtable_bemret_matlab.csv"), delimiter=",", skiprows=1)
    now = np.load(os.path.join(GOLD, "rod_quad_spectrum.npz"))
    rod = m.trirod(10.0, 40.0, n=[15, 15, 15])
    assert int(np.sum(rod.nvert == 4)) == 560
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsTable("gold.dat")], [rod], [[2, 1]], 1, interp="curv")
    pw = m.PlaneWaveRet([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    print("quadrilateral rod", precision, cp.n, flush=True)
    bem = m.BEMRet(cp, precision=precision)
    ext, sca = bem.solve_sweep(ml[:, 0], pw, chunk=41, with_sca=True)
    err = {"ext_x": np.max(np.abs(ext[:, 0] / ml[:, 1] - 1)), "sca_x": np.max(np.abs(sca[:, 0] / ml[:, 2] - 1)),
           "ext_z": np.max(np.abs(ext[:, 1] / ml[:, 3] - 1)), "sca_z": np.max(np.abs(sca[:, 1] / ml[:, 4] - 1))}
    print("max rel err vs the reference's spectra:", {k: "%.2e" % v for k, v in err.items()}, flush=True)
    assert max(err.values()) < tol, err
    idx = [int(np.argmin(np.abs(ml[:, 0] - lam))) for lam in now["enei"]]
    assert np.max(np.abs(ext[idx] / now["ext"] - 1)) < tol
    assert np.max(np.abs(sca[idx] / now["sca"] - 1)) < tol
