"""GPU check of the hipGraph-captured sweep step (``SweepGraph``): every
replay, with other eager work and synchronisations between replays and new
wavelengths written into the step's input buffers each time, against the eager
mixed-precision sweep of the same wavelengths (1e-10 relative; both refine to
1e-12 in FP64, the graph with a fixed number of steps)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DRUDE = (10.0, 9.065, 0.0708)


@pytest.mark.parametrize("residual", ["stored", "matrix_free"])
def test_sweep_graph_replays_match_eager(cuda, residual):
    import mnpbem_amd as m
    from mnpbem_amd.bem import SweepGraph
    cp = m.ComParticle([m.EpsConst(1.0), m.EpsDrude(*DRUDE)], [m.trisphere(144, 20.0)], [[2, 1]], 1)
    pw = m.PlaneWaveRet([[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]])
    bem = m.BEMRet(cp, precision="mixed", residual=residual)
    sets = [np.linspace(400.0, 800.0, 8), np.linspace(450.0, 700.0, 8), np.linspace(500.0, 520.0, 8)]
    eager = [bem.solve_sweep(l, pw, chunk=8) for l in sets]
    sg = SweepGraph(bem, pw, 8, iters=4)
    for rep, k in enumerate((0, 1, 2, 0, 1)):
        print("graph replay", residual, rep, flush=True)
        ext = sg.run(sets[k])
        assert np.abs(ext / eager[k] - 1).max() < 1e-10, (rep, k)
        assert sg.last_residual.max() <= bem.refine_tol
        # unrelated eager work between replays
        other = m.BEMRet(cp, precision="double").solve_sweep([610.0], pw)
        torch.cuda.synchronize()
        assert np.isfinite(other).all()
    assert sg.captures == 1
