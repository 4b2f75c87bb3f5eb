"""Quadrature rules on the unit triangle and their shape-function tables.

* ``rule18``: the 28-point, degree-11 rule of triangle_unit_set (published
  table; mnpbem/utils/quadrature.py:182-224), weights summing to 1.
* ``polar_rule``: polar integration about the centroid, three rotated
  60-degree sectors with Legendre-Gauss-Lobatto nodes in radius and angle
  (mnpbem/utils/quadface.py:106-182).
* ``tri6_tables``: quadratic shape functions and their x/y derivatives at the
  nodes, packed as the 19-column rows the GPU refinement kernel reads.
* ``polar_rule_quad`` / ``quad9_tables``: the polar rule of quadrilaterals
  (four rotated 45-degree sectors on [-1, 1]^2, weights summing to 4;
  mnpbem/utils/quadface.py:184-249) and the biquadratic 9-node shape
  functions (particle.py:1684-1750), packed as 28-column rows.
"""

import numpy as np


def lgl(n):
    """n+1 Legendre-Gauss-Lobatto nodes/weights on [-1, 1] (Newton iteration
    on x P_n - P_{n-1} from the Chebyshev-Lobatto points)."""
    x = np.cos(np.pi * np.arange(n + 1) / n)
    for _ in range(100):
        P = [np.ones_like(x), x.copy()]
        for k in range(2, n + 1):
            P.append(((2 * k - 1) * x * P[-1] - (k - 1) * P[-2]) / k)
        dx = (x * P[n] - P[n - 1]) / ((n + 1) * P[n])
        x = x - dx
        if np.max(np.abs(dx)) <= np.finfo(float).eps:
            break
    P = [np.ones_like(x), x.copy()]
    for k in range(2, n + 1):
        P.append(((2 * k - 1) * x * P[-1] - (k - 1) * P[-2]) / k)
    return x, 2.0 / (n * (n + 1) * P[n] ** 2)


_R18_A = [1.0 / 3.0, 0.9480217181434233, 0.02598914092828833, 0.8114249947041546,
          0.09428750264792270, 0.01072644996557060, 0.4946367750172147, 0.5853132347709715,
          0.2073433826145142, 0.1221843885990187, 0.4389078057004907, 0.6779376548825902,
          0.04484167758913055, 0.27722066752827925, 0.8588702812826364, 0.0, 0.1411297187173636]
_R18_W = [0.08797730116222190, 0.008744311553736190, 0.03808157199393533, 0.01885544805613125,
          0.07215969754474100, 0.06932913870553720, 0.04105631542928860, 0.007362383783300573]
# (x index, y index, weight index) into the tables above
_R18_PTS = [(0, 0, 0), (1, 2, 1), (2, 1, 1), (2, 2, 1), (3, 4, 2), (4, 3, 2), (4, 4, 2),
            (5, 6, 3), (6, 5, 3), (6, 6, 3), (7, 8, 4), (8, 7, 4), (8, 8, 4),
            (9, 10, 5), (10, 9, 5), (10, 10, 5), (11, 12, 6), (11, 13, 6), (12, 11, 6),
            (12, 13, 6), (13, 11, 6), (13, 12, 6), (14, 15, 7), (14, 16, 7), (15, 14, 7),
            (15, 16, 7), (16, 14, 7), (16, 15, 7)]


def rule18():
    x = np.array([_R18_A[a] for a, _, _ in _R18_PTS])
    y = np.array([_R18_A[b] for _, b, _ in _R18_PTS])
    w = np.array([_R18_W[c] for _, _, c in _R18_PTS])
    return x, y, w / w.sum()


def polar_rule(npol=(7, 5)):
    nr, na = (npol, npol) if np.isscalar(npol) else npol
    xr, wr = lgl(nr)
    xa, wa = lgl(na)
    rho = 0.5 * (xr + 1.0 + 1e-6)
    phi = np.deg2rad(270.0 + 60.0 * xa)
    rho2 = np.repeat(rho, len(phi))
    phi2 = np.tile(phi, len(rho))
    w2 = np.outer(wr, wa).ravel()
    edge = 1.0 / np.abs(2.0 * np.sin(phi2))
    xs, ys, ws = [], [], []
    for s in range(3):
        ang = phi2 + s * np.deg2rad(120.0)
        u = np.cos(ang) * rho2 * edge
        v = np.sin(ang) * rho2 * edge
        xs.append((1.0 - np.sqrt(3.0) * u - v) / 3.0)
        ys.append((1.0 + np.sqrt(3.0) * u - v) / 3.0)
        ws.append(w2 * rho2 * edge ** 2)
    w = np.concatenate(ws)
    return np.concatenate(xs), np.concatenate(ys), w / w.sum()


def tri6_tables(x, y, w):
    """Shape functions S (m, 6), derivatives Dx, Dy (m, 6) of the quadratic
    triangle with nodes (v0, v1, v2, m01, m12, m20) at barycentric
    (x, y, 1-x-y), and the weights."""
    l1, l2, l3 = x, y, 1.0 - x - y
    S = np.stack([l1 * (2 * l1 - 1), l2 * (2 * l2 - 1), l3 * (2 * l3 - 1),
                  4 * l1 * l2, 4 * l2 * l3, 4 * l3 * l1], axis=1)
    Dx = np.stack([4 * l1 - 1, 0 * x, 1 - 4 * l3, 4 * l2, -4 * l2, 4 * (l3 - l1)], axis=1)
    Dy = np.stack([0 * x, 4 * l2 - 1, 1 - 4 * l3, 4 * l1, 4 * (l3 - l2), -4 * l1], axis=1)
    return S, Dx, Dy, np.asarray(w, float)


def kernel_table(x, y, w):
    """(m, 19) rows [S(6), Dx(6), Dy(6), w] for the GPU kernel."""
    S, Dx, Dy, w = tri6_tables(x, y, w)
    return np.ascontiguousarray(np.hstack([S, Dx, Dy, w[:, None]]))


def polar_rule_quad(npol=(7, 5)):
    """Polar rule about the centre of [-1, 1]^2: LGL nodes in radius (mapped
    to (0, 1]) and angle (one 45-degree sector at 90-135 degrees, rotated
    four times by 90 degrees), weight w_rho w_phi rho r_edge^2 with r_edge =
    1 / |sin phi|, normalised to the square's area 4."""
    nr, na = (npol, npol) if np.isscalar(npol) else npol
    xr, wr = lgl(nr)
    xa, wa = lgl(na)
    rho = 0.5 * (xr + 1.0 + 1e-6)
    phi = np.deg2rad(90.0 + 45.0 * xa)
    rho2 = np.repeat(rho, len(phi))
    phi2 = np.tile(phi, len(rho))
    w2 = np.outer(wr, wa).ravel()
    edge = 1.0 / np.abs(np.sin(phi2))
    xs, ys, ws = [], [], []
    for s in range(4):
        ang = phi2 + s * np.deg2rad(90.0)
        xs.append(np.cos(ang) * rho2 * edge)
        ys.append(np.sin(ang) * rho2 * edge)
        ws.append(w2 * rho2 * edge ** 2)
    w = np.concatenate(ws)
    return np.concatenate(xs), np.concatenate(ys), 4.0 * w / w.sum()


def quad9_tables(x, y, w):
    """Biquadratic shape functions S (m, 9) and derivatives Dx, Dy of the
    9-node quadrilateral on [-1, 1]^2 with nodes (v0, v1, v2, v3) at
    (-1,-1), (1,-1), (1,1), (-1,1), edge midpoints (m01, m12, m23, m30) and
    the centre."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    S = np.stack([0.25 * x * (x - 1) * y * (y - 1), 0.25 * x * (x + 1) * y * (y - 1),
                  0.25 * x * (x + 1) * y * (y + 1), 0.25 * x * (x - 1) * y * (y + 1),
                  0.5 * (1 - x ** 2) * y * (y - 1), 0.5 * x * (x + 1) * (1 - y ** 2),
                  0.5 * (1 - x ** 2) * y * (y + 1), 0.5 * x * (x - 1) * (1 - y ** 2),
                  (1 - x ** 2) * (1 - y ** 2)], axis=1)
    Dx = np.stack([0.25 * (2 * x - 1) * y * (y - 1), 0.25 * (2 * x + 1) * y * (y - 1),
                   0.25 * (2 * x + 1) * y * (y + 1), 0.25 * (2 * x - 1) * y * (y + 1),
                   -x * y * (y - 1), 0.5 * (2 * x + 1) * (1 - y ** 2), -x * y * (y + 1),
                   0.5 * (2 * x - 1) * (1 - y ** 2), -2 * x * (1 - y ** 2)], axis=1)
    Dy = np.stack([0.25 * x * (x - 1) * (2 * y - 1), 0.25 * x * (x + 1) * (2 * y - 1),
                   0.25 * x * (x + 1) * (2 * y + 1), 0.25 * x * (x - 1) * (2 * y + 1),
                   0.5 * (1 - x ** 2) * (2 * y - 1), -x * (x + 1) * y, 0.5 * (1 - x ** 2) * (2 * y + 1),
                   -x * (x - 1) * y, -2 * (1 - x ** 2) * y], axis=1)
    return S, Dx, Dy, np.asarray(w, float)


def kernel_table9(x, y, w):
    """(m, 28) rows [S(9), Dx(9), Dy(9), w] of the quadrilateral polar rule."""
    S, Dx, Dy, w = quad9_tables(x, y, w)
    return np.ascontiguousarray(np.hstack([S, Dx, Dy, w[:, None]]))
