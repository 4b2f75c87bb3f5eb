"""CPU pins of the iterative-solve oracle (no GPU): the restated scipy GMRES
path of BEMRetIter reaches its tolerance and agrees with the direct solve,
for one and for two polarisations (as the reference's tests/test_bem_iter.py
and tests/test_iter_multipol.py assert), and the column-major packing
round-trips."""

import numpy as np
import pytest

import oracle as o
from oracle.iter_oracle import BEMRetIterOracle, pack, unpack


@pytest.fixture(scope="module")
def setup60():
    nodes, tri6 = o.trisphere_mesh(32, 10.0)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(10.0, 9.065, 0.0708)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    return o.GreenSetup(comp)


def test_pack_roundtrip():
    rng = np.random.RandomState(0)
    n, P = 7, 2
    parts = [rng.randn(n, P), rng.randn(n, 3, P), rng.randn(n, P), rng.randn(n, 3, P)]
    out = unpack(pack(*parts), n)
    for a, b in zip(parts, out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precond", [True, False])
def test_gmres_matches_direct(setup60, precond):
    comp = setup60.comp
    exc = o.planewave_potential(comp, [[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]], 550.0)
    d = o.BEMRetOracle(setup60).solve(exc)
    it = BEMRetIterOracle(setup60, tol=1e-10, maxit=400, precond=precond).solve(exc)
    assert it["flag"] == 0 and it["relres"] < 1e-9
    for f in ("sig1", "sig2", "h1", "h2"):
        rel = np.abs(it[f] - d[f]).max() / np.abs(d[f]).max()
        assert rel < (1e-9 if precond else 1e-6), f
