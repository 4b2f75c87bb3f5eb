"""Dielectric functions eps(lambda) -> (eps, k) (host scalars).

Same call contract as mnpbem/materials: ``eps_fun(enei)`` with the vacuum
wavelength in nm returns the complex permittivity and the wavenumber
k = 2 pi / lambda * sqrt(eps) in the material.
"""

import os

import numpy as np
from scipy.interpolate import CubicSpline

# photon energy (eV) <-> wavelength (nm), mnpbem/utils/constants.py:9
EV2NM = 1.0 / 8.0655477e-4

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def _k(enei, eps):
    return 2.0 * np.pi / enei * np.sqrt(eps)


class EpsConst:
    """Constant permittivity (mnpbem/materials/eps_const.py:48-72)."""

    def __init__(self, eps):
        if eps is None:
            raise ValueError("EpsConst: 'eps' must be a numeric value, got None.")
        if not isinstance(eps, (int, float, complex, np.integer, np.floating, np.complexfloating)):
            raise TypeError("EpsConst: 'eps' must be a numeric (int/float/complex) value, got %r."
                            % type(eps).__name__)
        self.eps = eps

    def __call__(self, enei):
        enei = np.asarray(enei, dtype=float)
        eps = np.full(enei.shape, self.eps, dtype=complex)
        out = (eps, _k(enei, eps))
        return (complex(out[0]), complex(out[1])) if enei.ndim == 0 else out


class EpsDrude:
    """Drude metal eps = eps0 - wp^2 / (w (w + i gammad)), w = EV2NM / lambda
    in eV (mnpbem/materials/eps_drude.py:73-107)."""

    def __init__(self, eps0, wp, gammad, name=None):
        self.eps0, self.wp, self.gammad, self.name = eps0, wp, gammad, name

    def __call__(self, enei):
        enei = np.asarray(enei, dtype=float)
        w = EV2NM / enei
        eps = self.eps0 - self.wp ** 2 / (w * (w + 1j * self.gammad))
        k = _k(enei, eps)
        return (complex(eps), complex(k)) if enei.ndim == 0 else (eps, k)

    def deriv(self, enei):
        """d eps / d(eps0, wp, gammad) at the wavelengths ``enei``, shape (..., 3)."""
        enei = np.asarray(enei, dtype=float)
        w = EV2NM / enei
        den = w * (w + 1j * self.gammad)
        d0 = np.ones(enei.shape, dtype=complex)
        dwp = -2.0 * self.wp / den
        dg = 1j * self.wp ** 2 / (w * (w + 1j * self.gammad) ** 2)
        return np.stack([d0, dwp, dg], axis=-1)


class EpsTable:
    """Tabulated refractive index (energy eV, n, k) with not-a-knot cubic
    splines of n and k over wavelength (mnpbem/materials/eps_table.py:51-152)."""

    def __init__(self, filename):
        path = filename if os.path.exists(filename) else os.path.join(_DATA, filename)
        if not os.path.exists(path):
            raise FileNotFoundError("Material data file not found: %s" % filename)
        rows = []
        with open(path) as fh:
            for line in fh:
                s = line.strip()
                if not s or s.startswith("%") or s.startswith("#"):
                    continue
                try:
                    v = [float(x) for x in s.split()]
                except ValueError:
                    continue
                if len(v) >= 3:
                    rows.append(v[:3])
        if not rows:
            raise ValueError("No valid data found in %s" % path)
        d = np.array(rows)
        lam = EV2NM / d[:, 0]
        o = np.argsort(lam)
        self.enei = lam[o]
        self.ni = CubicSpline(self.enei, d[o, 1])
        self.ki = CubicSpline(self.enei, d[o, 2])
        self.filename = os.path.basename(path)

    def __call__(self, enei):
        enei = np.asarray(enei, dtype=float)
        if np.any(enei < self.enei.min()) or np.any(enei > self.enei.max()):
            raise ValueError("Wavelength out of range. Valid range: %.1f - %.1f nm"
                             % (self.enei.min(), self.enei.max()))
        n = self.ni(enei) + 1j * self.ki(enei)
        eps = n * n
        k = _k(enei, eps)
        return (complex(eps), complex(k)) if enei.ndim == 0 else (eps, k)
