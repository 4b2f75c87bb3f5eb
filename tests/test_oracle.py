"""Pins of the FP64 CPU oracle (no GPU).

The oracle is checked against (a) the reference's committed spectra for
trisphere(144, 20) with tabulated gold (validation/03_bemret/sphere/data),
(b) closed forms and quadrature exactness, (c) brute-force integrals of the
refined Green-function entries, (d) the Fuchs-Liu identity in the static
limit and (e) the 8N BEM operator identity of the direct solve.
"""

import math
import os

import numpy as np
import pytest

import oracle as o

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def sphere284():
    nodes, tri6 = o.trisphere_mesh(144, 20.0)
    return o.CurvedSurface(nodes, tri6)


def test_golden_spectrum_tabulated_gold(sphere284):
    """Whole path vs the reference's committed BEMRet spectra (printed with 11
    significant digits): extinction vs the Python run, scattering vs MATLAB."""
    py = np.loadtxt(os.path.join(GOLD, "sphere144_d20_goldtable_bemret_python.csv"), delimiter=",", skiprows=1)
    ml = np.loadtxt(os.path.join(GOLD, "sphere144_d20_goldtable_bemret_matlab.csv"), delimiter=",", skiprows=1)
    comp = o.Composite([o.EpsConst(1.0), o.EpsTable(os.path.join(o.DATA_DIR, "gold.dat"))], [sphere284], [[2, 1]])
    sel = np.arange(0, 41, 4)
    ext, sca = o.solve_spectrum(comp, py[sel, 0], [[1, 0, 0]], [[0, 0, 1]], with_sca=True)
    assert np.max(np.abs(ext[:, 0] / py[sel, 1] - 1)) < 2e-10
    assert np.max(np.abs(ext[:, 0] / ml[sel, 1] - 1)) < 2e-10
    assert np.max(np.abs(sca[:, 0] / ml[sel, 2] - 1)) < 2e-10
    # absorption = ext - sca is positive for a lossy metal
    assert np.all(ext[:, 0] - sca[:, 0] > 0)


def test_drude_closed_form():
    d = o.EpsDrude(10.0, 9.065, 0.0708)
    for lam in (400.0, 520.0, 800.0):
        w = (1.0 / 8.0655477e-4) / lam  # photon energy in eV, conversion of utils/constants.py:9
        eps_ref = 10.0 - 9.065 ** 2 / (w * (w + 0.0708j))
        eps, k = d(lam)
        assert abs(eps - eps_ref) < 1e-14 * abs(eps_ref)
        assert abs(k - 2 * np.pi / lam * np.sqrt(eps)) < 1e-15


def test_ev2nm_is_hc():
    assert abs(o.EV2NM / 1239.8419843320026 - 1) < 1e-6  # hc in eV nm (CODATA 2018)


def test_table_interpolates_knots():
    t = o.EpsTable(os.path.join(o.DATA_DIR, "gold.dat"))
    rows = np.array([[float(x) for x in l.split()] for l in open(os.path.join(o.DATA_DIR, "gold.dat"))
                     if l.strip() and l.strip()[0] not in "%#"])
    for e, n, k in rows[5:40:7]:
        eps, _ = t(o.EV2NM / e)
        assert abs(eps - (n + 1j * k) ** 2) < 1e-12 * abs(eps)


def test_rule18_exact_to_degree_11():
    x, y, w = o.rule18()
    assert len(w) == 28 and abs(w.sum() - 1) < 1e-15
    for a in range(12):
        for b in range(12 - a):
            exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(w * x ** a * y ** b) - exact) < 1e-14


def test_lgl_exactness():
    x, w = o.lglnodes(7)
    assert len(x) == 8 and abs(w.sum() - 2) < 1e-14
    for p in range(14):  # exact for degree 2n-1 = 13
        assert abs(np.sum(w * x ** p) - (0 if p % 2 else 2.0 / (p + 1))) < 1e-13


def test_polar_rule_integrates_smooth_functions():
    x, y, w = o.polar_rule_tri((7, 5))
    assert len(w) == 3 * 8 * 6 and abs(w.sum() - 1) < 1e-14
    for a, b in [(1, 0), (0, 1), (2, 1), (3, 3)]:
        exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
        assert abs(np.sum(w * x ** a * y ** b) - exact) < 5e-3 * exact  # not exact: built for 1/r


def test_curved_sphere_geometry(sphere284):
    s = sphere284
    assert s.n == 284
    assert np.allclose(np.linalg.norm(s.nvec, axis=1), 1, atol=1e-14)
    assert np.all(np.sum(s.nvec * s.pos, axis=1) > 0)  # outward
    assert abs(s.area.sum() / (np.pi * 400.0) - 1) < 2e-4
    nodes, tri6 = o.trisphere_mesh(1444, 20.0)
    big = o.CurvedSurface(nodes, tri6)
    assert big.n == 2884
    assert abs(big.area.sum() / (np.pi * 400.0) - 1) < 2e-5


def _brute_green(surf, i, j, k, sub=10):
    """G_ij and F_ij by a subdivided 28-point rule on the curved face j."""
    x, y, w = o.rule18()
    X, Y, Wt = [], [], []
    h = 1.0 / sub
    for a in range(sub):
        for b in range(sub - a):
            X.append((a + x) * h); Y.append((b + y) * h); Wt.append(w)
            if b != sub - 1 - a:
                X.append((a + 1 - x) * h); Y.append((b + 1 - y) * h); Wt.append(w)
    X, Y, Wt = np.concatenate(X), np.concatenate(Y), np.concatenate(Wt)
    Wt = Wt / Wt.sum()
    Xn = surf.nodes[surf.tri6[j]]
    S = o.tri6_shape(X, Y)
    Dx, Dy = o.tri6_deriv(X, Y)
    pts = S @ Xn
    jac = np.linalg.norm(np.cross(Dx @ Xn, Dy @ Xn), axis=1)
    wq = 0.5 * Wt * jac
    rv = surf.pos[i] - pts
    r = np.linalg.norm(rv, axis=1)
    G = np.sum(wq * np.exp(1j * k * r) / r)
    F = np.sum(wq * (rv @ surf.nvec[i]) * (1j * k - 1 / r) / r ** 2 * np.exp(1j * k * r))
    return G, F


def test_refined_entries_match_brute_force(sphere284):
    """Near-pair G and F from the order-5 expansion vs a 2800-point
    quadrature of the exact kernel (no singularity at near pairs)."""
    s = sphere284
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(10.0, 9.065, 0.0708)], [s], [[2, 1]])
    setup = o.GreenSetup(comp)
    lam = 520.0
    kin = comp.eps[1](lam)[1]
    G1, H1, G2, H2 = o.green_matrices(setup, lam)
    row, col, _, _ = setup.blocks[(0, 0)]
    near = np.nonzero(row != col)[0]
    rng = np.random.default_rng(3)
    for q in rng.choice(near, 6, replace=False):
        i, j = row[q], col[q]
        Gb, Fb = _brute_green(s, i, j, kin)
        assert abs(G1[i, j] - Gb) < 1e-7 * abs(Gb)
        assert abs(H1[i, j] - Fb) < 1e-5 * abs(Fb) + 1e-9
    # a far pair is the one-point rule: G = a_j e^{ikd}/d, far from the exact integral only by quadrature
    far = np.argwhere(o.refinematrix(s.pos, s.pos, s.bradius, s.bradius) == 0)[0]
    Gb, _ = _brute_green(s, far[0], far[1], kin)
    assert abs(G1[far[0], far[1]] - Gb) < 5e-3 * abs(Gb)


def test_fuchs_liu_static_limit(sphere284):
    """As k -> 0 the refined, corrected F satisfies sum_i a_i F_ij / a_j = -2 pi
    for every face j of a closed surface (Fuchs & Liu, PRB 14, 5521)."""
    s = sphere284
    comp = o.Composite([o.EpsConst(1.0), o.EpsConst(1.0)], [s], [[2, 1]])
    setup = o.GreenSetup(comp)
    _, H1, _, _ = o.green_matrices(setup, 1e9)
    F = H1 - 2 * np.pi * np.eye(s.n)
    colsum = (s.area[:, None] * F / s.area[None, :]).sum(axis=0)
    assert np.max(np.abs(colsum + 2 * np.pi)) < 1e-6


def test_direct_solve_satisfies_bem_operator(sphere284):
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(10.0, 9.065, 0.0708)], [sphere284], [[2, 1]])
    setup = o.GreenSetup(comp)
    bem = o.BEMRetOracle(setup)
    for lam in (450.0, 520.0):
        exc = o.planewave_potential(comp, [[1, 0, 0], [0, 1, 0]], [[0, 0, 1], [0, 0, 1]], lam)
        sig = bem.solve(exc)
        phi, a, alpha, De = o.bem_operator_apply(bem.mats, bem.eps1, bem.eps2, bem.k, comp.nvec, sig)
        a_ref = exc["a2"] - exc["a1"]
        al_ref = exc["a2p"] - exc["a1p"]
        De_ref = -1j * bem.k * (bem.eps2 * np.einsum("id,idp->ip", comp.nvec, exc["a2"])
                                - bem.eps1 * np.einsum("id,idp->ip", comp.nvec, exc["a1"]))
        scale = np.abs(al_ref).max()
        assert np.abs(phi).max() < 1e-11 * np.abs(a_ref).max()
        assert np.abs(a - a_ref).max() < 1e-11 * np.abs(a_ref).max()
        assert np.abs(alpha - al_ref).max() < 1e-10 * scale
        assert np.abs(De - De_ref).max() < 1e-10 * np.abs(De_ref).max()


def test_cartesian_green_derivative(sphere284):
    """Gp of the oracle: (a) the closed-form gradient of G against central
    differences of G in the field point at a far pair, (b) n . Gp equals F
    for every pair, refined ones included (F is pinned by brute force)."""
    comp = o.Composite([o.EpsConst(1.0), o.EpsDrude(10.0, 9.065, 0.0708)], [sphere284], [[2, 1]])
    setup = o.GreenSetup(comp)
    lam = 520.0
    Hp = o.green_hp(setup, lam, 0)
    G1, H1, _, _ = o.green_matrices(setup, lam)
    k = comp.eps[1](lam)[1]
    n = comp.n
    F = H1 - 2 * np.pi * np.eye(n)
    nGp = np.einsum("ic,icj->ij", sphere284.nvec, Hp) - 2 * np.pi * np.eye(n)
    assert np.abs(nGp - F).max() < 1e-13 * np.abs(F).max()
    i, j = 3, 200
    ir = o.refinematrix(sphere284.pos, sphere284.pos, sphere284.bradius, sphere284.bradius)
    assert ir[i, j] == 0
    gfun = lambda r: sphere284.area[j] * np.exp(1j * k * np.linalg.norm(r - sphere284.pos[j])) / np.linalg.norm(
        r - sphere284.pos[j])
    hstep = 1e-5
    num = [(gfun(sphere284.pos[i] + hstep * e) - gfun(sphere284.pos[i] - hstep * e)) / (2 * hstep)
           for e in np.eye(3)]
    assert np.allclose(num, Hp[i, :, j], rtol=1e-7, atol=0)
