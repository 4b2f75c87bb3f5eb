"""GPU checks of the adjoint (transposed) BEM solve A^T y = c that the
vertex-position gradient rests on, through the C ABI (via the Python API).

The transposed operator is pinned against the forward one by the bilinear
identity y . (A x) = (A^T y) . x on random complex vectors (the forward
operator is pinned against the oracle elsewhere), and the adjoint solve by
A^T y = c and y . r = c . x for A x = r.  Tolerances: complex128 path
1e-11 relative, mixed path (complex64 factors + FP64 refinement) 1e-10.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DRUDE = (10.0, 9.065, 0.0708)


def _rand(shape, seed, dev):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return torch.as_tensor(a, dtype=torch.complex128).to(dev)


def _dot(a, b):
    return (a * b).sum(dim=(1, 2))


def _particles(m):
    yield "sphere284", m.trisphere(144, 20.0)
    yield "rod", m.trirod(10.0, 24.0, [10, 5, 6], triangles=True)


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_adjoint_solve(cuda, precision, inverse):
    """The forward and adjoint solves on the LU factors and, with
    ``inverse``, on the explicit inverses that BEMRet.invert_factors leaves in
    their place (mnpbem_bem_solve_inv_batched,
    mnpbem_bem_solve_adjoint_inv_batched): same checks, same tolerances."""
    import mnpbem_amd as m
    lams = np.array([480.0, 560.0, 700.0])
    for name, part in _particles(m):
        print("adjoint solve", name, precision, "inverse" if inverse else "factors", part.n, flush=True)
        cp = m.ComParticle([m.EpsConst(1.33 ** 2), m.EpsDrude(*DRUDE)], [part], [[2, 1]], 1)
        bem = m.BEMRet(cp, precision=precision)
        F = bem.factor(lams)
        if inverse:
            bem.invert_factors(F)
        n, B = cp.n, len(lams)
        for P in (1, 2):
            X = _rand((B, n, 8 * P), 1, cuda)
            Yr = _rand((B, n, 8 * P), 2, cuda)
            # transposed operator vs forward operator
            lhs = _dot(Yr, bem.apply(F, X))
            rhs = _dot(bem.apply_transpose(F, Yr), X)
            assert torch.max(torch.abs(lhs - rhs) / torch.abs(lhs)).item() < 1e-12, (name, P)
            # adjoint solve
            C = _rand((B, n, 8 * P), 3, cuda)
            Y = bem.solve_adjoint_batch(F, C)
            res = bem.apply_transpose(F, Y) - C
            rel = torch.linalg.vector_norm(res, dim=(1, 2)) / torch.linalg.vector_norm(C, dim=(1, 2))
            tol = 1e-11 if precision == "double" else 1e-10
            assert rel.max().item() < tol, (name, P, rel)
            R = _rand((B, n, 8 * P), 4, cuda)
            Xs = bem.solve_batch(F, R)
            a, b = _dot(Y, R), _dot(C, Xs)
            assert torch.max(torch.abs(a - b) / torch.abs(b)).item() < tol, (name, P)


def test_material_change_refactors(cuda):
    """A solver whose factors are cached for a wavelength factors again when
    a material is changed in place (as DrudeExtinction does with the new
    parameters), so bem.solve at that wavelength sees the new material."""
    import mnpbem_amd as m
    from mnpbem_amd.adjoint import DrudeExtinction
    drude = m.EpsDrude(*DRUDE)
    cp = m.ComParticle([m.EpsConst(1.0), drude], [m.trisphere(60, 20.0)], [[2, 1]], 1)
    pw = m.PlaneWaveRet([[1, 0, 0]], [[0, 0, 1]])
    bem = m.BEMRet(cp)
    lam = 540.0
    print("material change, cached factors", flush=True)
    e_old = pw.extinction(bem.solve(pw.potential(cp, lam))[0])
    params = torch.tensor([9.0, 8.9, 0.09], dtype=torch.float64, requires_grad=True)
    DrudeExtinction.apply(params, bem, [lam], pw)
    e_new = pw.extinction(bem.solve(pw.potential(cp, lam))[0])
    fresh = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(9.0, 8.9, 0.09)], [m.trisphere(60, 20.0)], [[2, 1]], 1)
    e_fresh = pw.extinction(m.BEMRet(fresh).solve(pw.potential(fresh, lam))[0])
    assert abs(e_new / e_fresh - 1) < 1e-12
    assert abs(e_new / e_old - 1) > 1e-3
    # an in-place change without the wrapper is caught as well
    drude.gammad = 0.2
    e_g = pw.extinction(bem.solve(pw.potential(cp, lam))[0])
    fresh2 = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(9.0, 8.9, 0.2)], [m.trisphere(60, 20.0)], [[2, 1]], 1)
    assert abs(e_g / pw.extinction(m.BEMRet(fresh2).solve(pw.potential(fresh2, lam))[0]) - 1) < 1e-12
