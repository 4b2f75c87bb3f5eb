"""CPU checks of the stopping rule of the mixed-precision FP64 refinement
(``mnpbem_amd.bem.RefinementMonitor``): it stops when every system meets the
tolerance, keeps each system's best iterate, stops on stagnation or
divergence and then raises instead of returning an unconverged solution.

The numeric cases drive the monitor with the same loop the GPU path runs
(low-precision solve, FP64 residual, correction) written in numpy: a
well-conditioned system converges; a nearly singular one (condition 1e9, far
beyond what complex64 factors can refine) stagnates and raises."""

import numpy as np
import pytest

from mnpbem_amd.bem import RefinementMonitor


def _refine(A, b, tol, maxit=8):
    """x <- x + solve32(b - A x) with the monitor's stopping rule; returns the
    best iterate and the monitor."""
    A32 = A.astype(np.complex64)
    mon = RefinementMonitor(tol, maxit)
    x = np.zeros_like(b)
    best = None
    r = b.copy()
    while True:
        x = x + np.linalg.solve(A32, r.astype(np.complex64)).astype(np.complex128)
        r = b - A @ x
        rel = np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)
        improved, done = mon.step(rel)
        best = x.copy() if best is None else np.where(improved[None, :], x, best)
        if done:
            return best, mon


def _matrix(n, cond, seed):
    rng = np.random.default_rng(seed)
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    s = np.geomspace(1.0, 1.0 / cond, n)
    return (q1 * s) @ q2


def test_well_conditioned_system_converges():
    A = _matrix(80, 1e2, 1)
    b = np.random.default_rng(2).standard_normal((80, 3)) + 0j
    x, mon = _refine(A, b, 1e-12)
    mon.check("test")
    assert np.all(mon.best <= 1e-12)
    assert 2 <= mon.it <= 5
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-12


def test_nearly_singular_system_raises():
    A = _matrix(80, 1e9, 3)
    b = np.random.default_rng(4).standard_normal((80, 2)) + 0j
    x, mon = _refine(A, b, 1e-12)
    assert np.any(mon.best > 1e-12)
    assert mon.it <= 9
    with pytest.raises(RuntimeError, match="refine_tol"):
        mon.check("BEMRet.solve")
    # the returned iterate is the best one seen, not the last
    rel = np.linalg.norm(A @ x - b, axis=0) / np.linalg.norm(b, axis=0)
    assert np.allclose(rel, mon.best, rtol=1e-6)


def test_sequences():
    # converges in three steps
    m = RefinementMonitor(1e-12, 8)
    imp, done = m.step([1e-4, 1e-5])
    assert imp.all() and not done
    imp, done = m.step([1e-8, 1e-9])
    assert imp.all() and not done
    imp, done = m.step([1e-13, 1e-13])
    assert done
    m.check("x")
    # divergence: the second system gets worse, the first is already converged
    m = RefinementMonitor(1e-12, 8)
    m.step([1e-13, 1e-6])
    imp, done = m.step([1e-14, 1e-3])
    assert list(imp) == [True, False] and done
    assert m.best[1] == 1e-6
    with pytest.raises(RuntimeError):
        m.check("x")
    # slow but steady progress continues up to the step limit
    m = RefinementMonitor(1e-12, 3)
    for k, r in enumerate([1e-2, 4e-3, 1.6e-3, 6.4e-4]):
        imp, done = m.step([r])
        assert done == (k == 3)
    with pytest.raises(RuntimeError):
        m.check("x")
