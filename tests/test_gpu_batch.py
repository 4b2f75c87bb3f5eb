"""GPU parity of the batched multi-particle solve (BEMRetBatch: particles
grouped by face count, each group's matrices for a wavelength block in one
batch with per-particle geometry slots) against the per-particle BEMRet
solves and the CPU oracle, through the C ABI.

Six seeded Drude-gold particles (spheres with curved faces and flat
ellipsoids of 284, 508 and 796 faces, gamma U(0.05, 0.1) eV), two plane
waves, three wavelengths.  Bars: complex128 batched vs per particle 1e-12
(same kernels; only the triangular-leaf choice, which depends on the batch's
column count, can move rounding); mixed batched vs per-particle mixed 1e-10
and vs per-particle complex128 1e-9; the two smallest particles against the
FP64 oracle 1e-10 (complex128) / 1e-9 (mixed); the captured block replayed
as a HIP graph against the eager batched solve 1e-12 (fixed refinement
steps instead of the stopping rule).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle as o  # noqa: E402

LAMS = np.array([450.0, 560.0, 690.0])
POL2 = ([[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]])


def _specs():
    rng = np.random.default_rng(3)
    kinds = [("sphere", 144), ("ellipsoid", 144), ("sphere", 256), ("ellipsoid", 400), ("sphere", 144),
             ("ellipsoid", 256)]
    out = []
    for kind, nv in kinds:
        geo = float(rng.uniform(20.0, 40.0)) if kind == "sphere" else rng.uniform(10.0, 20.0, 3)
        out.append((kind, nv, geo, float(rng.uniform(0.05, 0.1))))
    return out


def _particle(m, spec):
    kind, nv, geo, gam = spec
    part = m.trisphere(nv, geo) if kind == "sphere" else m.triellipsoid(nv, geo)
    return m.ComParticle([m.EpsConst(1.0), m.EpsDrude(10.0, 9.065, gam)], [part], [[2, 1]], 1)


def _oracle(spec):
    kind, nv, geo, gam = spec
    nodes, tri6 = o.trisphere_mesh(nv, geo) if kind == "sphere" else o.triellipsoid_mesh(nv, geo)
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(10.0, 9.065, gam)], [o.CurvedSurface(nodes, tri6)], [[2, 1]])
    return o.solve_spectrum(comp, LAMS, *POL2)


@pytest.fixture(scope="module")
def batch_setup(cuda):
    import mnpbem_amd as m
    specs = _specs()
    parts = [_particle(m, s) for s in specs]
    pw = m.PlaneWaveRet(*POL2)
    print("per-particle reference solves", [p.n for p in parts], flush=True)
    single = {prec: [m.BEMRet(p, precision=prec).solve_sweep(LAMS, pw) for p in parts]
              for prec in ("double", "mixed")}
    return specs, parts, pw, single


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_batched_matches_per_particle(batch_setup, precision):
    from mnpbem_amd.batch import BEMRetBatch
    specs, parts, pw, single = batch_setup
    print("BEMRetBatch", precision, flush=True)
    bb = BEMRetBatch(parts, precision=precision)
    assert sorted(g.n for g in bb.groups) == [284, 508, 796]
    assert sorted(g.S for g in bb.groups) == [1, 2, 3]
    ext = bb.solve_sweep(LAMS, pw, block=3)
    for k, p in enumerate(parts):
        same = np.max(np.abs(ext[k] / single[precision][k] - 1))
        print("particle %d (%d faces): batched vs per-particle %.2e" % (k, p.n, same), flush=True)
        assert same < (1e-12 if precision == "double" else 1e-10), (k, same)
        assert np.max(np.abs(ext[k] / single["double"][k] - 1)) < 1e-9
    if precision == "mixed":
        assert bb.last_residual <= bb.refine_tol


def test_batched_smallest_against_oracle(batch_setup):
    from mnpbem_amd.batch import BEMRetBatch
    specs, parts, pw, single = batch_setup
    ext = {prec: BEMRetBatch(parts, precision=prec).solve_sweep(LAMS, pw, block=3) for prec in ("double", "mixed")}
    for k in (0, 1):   # the 284-face sphere and ellipsoid
        print("oracle, particle", k, specs[k][0], flush=True)
        ref = _oracle(specs[k])
        assert np.max(np.abs(ext["double"][k] / ref - 1)) < 1e-10
        assert np.max(np.abs(ext["mixed"][k] / ref - 1)) < 1e-9


def test_batched_graph_replay(batch_setup):
    from mnpbem_amd.batch import BEMRetBatch, BatchSweepGraph
    specs, parts, pw, single = batch_setup
    bb = BEMRetBatch(parts, precision="mixed")
    eager = bb.solve_sweep(LAMS, pw, block=3)
    print("BatchSweepGraph L=3", flush=True)
    sg = BatchSweepGraph(bb, pw, 3, iters=3)
    first = sg.solve_sweep(LAMS)
    bb.solve_sweep(LAMS[:1], pw, block=1)           # eager work between replays
    second = sg.solve_sweep(LAMS)
    for k in range(len(parts)):
        assert np.array_equal(first[k], second[k])
        assert np.max(np.abs(second[k] / eager[k] - 1)) < 1e-12
    assert sg.last_residual <= bb.refine_tol


def test_batched_graph_too_few_steps_resolves_groups(batch_setup):
    """A replay whose captured refinement steps miss refine_tol (one step
    captured) solves the failing groups again eagerly, group by group, and
    keeps the replay's results for the rest: the spectra equal the eager
    solve's and the final residual meets refine_tol."""
    from mnpbem_amd.batch import BEMRetBatch, BatchSweepGraph
    specs, parts, pw, single = batch_setup
    bb = BEMRetBatch(parts, precision="mixed")
    eager = bb.solve_sweep(LAMS, pw, block=3)
    print("BatchSweepGraph L=3 iters=1", flush=True)
    sg = BatchSweepGraph(bb, pw, 3, iters=1)
    out = sg.solve_sweep(LAMS)
    assert sg.fallbacks > 0 and sg.last_residual <= bb.refine_tol
    for k in range(len(parts)):
        assert np.max(np.abs(out[k] / eager[k] - 1)) < 1e-12
