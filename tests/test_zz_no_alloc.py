"""GPU check of mnpbem_bem_solve_batched_noalloc (the BEM solve with its work
space in the library's static device array) against
mnpbem_bem_solve_batched (caller-supplied work space), through the C ABI:
two calls in a row, the second with a smaller input (fewer faces and
wavelengths, so stale contents of the shared array from the first call
would show), for the complex128 and the complex64 (bf16x3) factors.  Both
entry points run the same kernels in the same order, so the solutions are
bit-identical.  A batch whose work space exceeds the array is rejected on
the host before any launch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DRUDE = (10.0, 9.065, 0.0708)


def _solve_both(m, nat, nv, lams, precision):
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(nv, 20.0)], [[2, 1]], 1)
    bem = m.BEMRet(cp, precision=precision)
    pw = m.PlaneWaveRet([[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]])
    F = bem.factor(lams)
    R = pw.rhs(cp, lams, bem.g)
    if F.dtype != nat.DTYPE_C128:
        R = R.to(torch.complex64)
    P = R.shape[2] // 8
    n, nn = F.n, F.n * F.n
    st = nat.stream_handle(bem.device)
    X1 = torch.empty_like(R)
    W = torch.empty_like(R)
    common = (F.dtype, n, P, nat.ptr(F.GG[0]), nat.ptr(F.piv[0]), nat.ptr(F.GG[1]), nat.ptr(F.piv[1]),
              nat.ptr(F.HH[0]), nat.ptr(F.Et), nat.ptr(F.HH[1]), nat.ptr(F.piv[3]), n, nn, n, nat.ptr(F.DG[0]),
              nat.ptr(F.DG[1]), nat.ptr(F.DS), F.sD, nat.ptr(R))
    tail = (8 * P, n * 8 * P, nat.ptr(bem.g.nvec), nat.ptr(F.par), 0, F.B, st)
    print("solve with caller work space: N=%d B=%d %s" % (n, F.B, precision), flush=True)
    nat.call("mnpbem_bem_solve_batched", *common, nat.ptr(W), nat.ptr(X1), *tail)
    X2 = torch.full_like(R, float("nan"))
    print("solve with static scratch: N=%d B=%d %s" % (n, F.B, precision), flush=True)
    nat.call("mnpbem_bem_solve_batched_noalloc", *common, nat.ptr(X2), *tail)
    torch.cuda.synchronize()
    return X1, X2


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_noalloc_solve_matches_workspace_solve(cuda, precision):
    import mnpbem_amd as m
    from mnpbem_amd import _native as nat
    first = _solve_both(m, nat, 256, np.array([450.0, 520.0, 610.0, 700.0]), precision)
    second = _solve_both(m, nat, 60, np.array([480.0, 590.0]), precision)
    for X1, X2 in (first, second):
        assert torch.isfinite(X2).all()
        assert torch.equal(X1, X2)


def test_noalloc_rejects_oversized_batch(cuda):
    """((batch-1) sv + (n-1) ldv + 8P) x 16 bytes above 64 MiB is refused
    before anything is launched (pointers are never dereferenced)."""
    from mnpbem_amd import _native as nat
    n, P, B = 2884, 2, 200
    need = ((B - 1) * n * 8 * P + (n - 1) * 8 * P + 8 * P) * 16
    assert need > 2 * (64 << 20)   # far above the array: the host check must refuse it
    with pytest.raises(RuntimeError, match="exceeds the static"):
        nat.call("mnpbem_bem_solve_batched_noalloc", nat.DTYPE_C128, n, P, None, None, None, None, None, None,
                 None, None, n, n * n, n, None, None, None, 0, None, None, 8 * P, n * 8 * P, None, None, 0, B,
                 nat.stream_handle(cuda))
