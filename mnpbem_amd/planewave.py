"""Plane-wave excitation and its extinction cross section
(mnpbem/simulation/planewave_ret.py, mnpbem/spectrum/spectrum_ret.py)."""

import numpy as np
import torch

from . import _native as nat
from .bem import CompStruct


class PlaneWaveRet:
    """Plane wave with polarisations ``pol`` (npol, 3) and propagation
    directions ``dir`` (npol, 3) in medium ``medium`` (1-based)."""

    def __init__(self, pol, dir, medium=1, **options):
        if pol is None:
            raise ValueError("PlaneWaveRet: 'pol' must be a polarization vector, got None.")
        if dir is None:
            raise ValueError("PlaneWaveRet: 'dir' must be a propagation direction vector, got None.")
        self.pol = np.atleast_2d(np.asarray(pol, dtype=float))
        self.dir = np.atleast_2d(np.asarray(dir, dtype=float))
        if self.pol.shape[1] != 3 or self.dir.shape[1] != 3 or self.pol.ndim != 2:
            raise ValueError("PlaneWaveRet: 'pol' and 'dir' must have shape (3,) or (npol, 3).")
        if self.pol.shape != self.dir.shape:
            raise ValueError("PlaneWaveRet: 'pol' and 'dir' must have the same number of rows.")
        if not np.allclose(np.sum(self.pol * self.dir, axis=1), 0):
            raise ValueError("Polarization and propagation direction must be orthogonal")
        self.medium = options.get("medium", medium)

    @property
    def npol(self):
        return self.pol.shape[0]

    def _pw(self, p, enei_list, device):
        """Per wavelength {k0, Re k, Im k, eps1, eps2} as a device tensor."""
        rows = []
        m_in, m_out = p.inout[0, 0], p.inout[0, 1]
        for lam in enei_list:
            k0 = 2.0 * np.pi / lam
            eps_m = complex(p.eps[self.medium - 1](lam)[0])
            k = k0 * np.sqrt(eps_m)
            e1 = complex(p.eps[m_in - 1](lam)[0])
            e2 = complex(p.eps[m_out - 1](lam)[0])
            rows.append([k0, k.real, k.imag, e1.real, e1.imag, e2.real, e2.imag])
        return torch.as_tensor(np.array(rows), dtype=torch.float64).to(device)

    def _side_flags(self, p, g):
        key = ("_flags", id(g))
        fl = getattr(self, "_flag_cache", None)
        if fl is None or fl[0] != key:
            inm = torch.as_tensor((p.inout_faces[:, 0] == self.medium).astype(np.int32)).to(g.device)
            outm = torch.as_tensor((p.inout_faces[:, 1] == self.medium).astype(np.int32)).to(g.device)
            pol = torch.as_tensor(self.pol, dtype=torch.float64).to(g.device)
            dr = torch.as_tensor(self.dir, dtype=torch.float64).to(g.device)
            self._flag_cache = (key, (inm, outm, pol, dr))
        return self._flag_cache[1]

    def rhs(self, p, enei_list, g, potentials=False, pw=None):
        """BEM right side R (B, n, 8P) = [phi | a | De | alpha] (bem_ret.py:1360-1429)
        and optionally the potentials (B, n, 12P) = [a1 | a1' | a2 | a2'].
        ``pw``: the per-wavelength parameters already on the device (``_pw``)."""
        enei_list = np.atleast_1d(enei_list)
        B, n, P = len(enei_list), p.n, self.npol
        inm, outm, pol, dr = self._side_flags(p, g)
        pw = self._pw(p, enei_list, g.device) if pw is None else pw
        R = torch.empty((B, n, 8 * P), dtype=torch.complex128, device=g.device)
        pot = torch.empty((B, n, 12 * P), dtype=torch.complex128, device=g.device) if potentials else None
        nat.call("mnpbem_planewave_batched", n, P, nat.ptr(g.pos), nat.ptr(g.nvec), nat.ptr(inm), nat.ptr(outm),
                 nat.ptr(pol), nat.ptr(dr), nat.ptr(pw), nat.ptr(R), 8 * P, n * 8 * P, nat.ptr(pot), n * 12 * P, 0,
                 B, nat.stream_handle(g.device))
        return (R, pot) if potentials else R

    def potential(self, p, enei, green=None):
        """CompStruct with a1, a1p, a2, a2p (n, 3, npol), squeezed for npol 1
        (planewave_ret.py:257-361).  ``green``: the CompGreenRet whose device
        geometry to use (built on demand)."""
        g = green if green is not None else _green_for(p)
        R, pot = self.rhs(p, [enei], g, potentials=True)
        n, P = p.n, self.npol
        f = {}
        for q, name in enumerate(("a1", "a1p", "a2", "a2p")):
            v = pot[0, :, 3 * P * q:3 * P * (q + 1)].reshape(n, 3, P)
            f[name] = v[:, :, 0] if P == 1 else v
        return CompStruct(p, enei, _R=R, **f)

    __call__ = potential

    def extinction_batch(self, p, enei_list, X, g, pw=None):
        """Extinction (B, npol) from solutions X (B, n, 8P)."""
        enei_list = np.atleast_1d(enei_list)
        B, n, P = len(enei_list), p.n, self.npol
        inm, outm, pol, dr = self._side_flags(p, g)
        pw = self._pw(p, enei_list, g.device) if pw is None else pw
        ext = torch.empty((B, P), dtype=torch.float64, device=g.device)
        nat.call("mnpbem_extinction_batched", n, P, nat.ptr(X), 8 * P, n * 8 * P, nat.ptr(g.pos), nat.ptr(g.area),
                 nat.ptr(inm), nat.ptr(outm), nat.ptr(pol), nat.ptr(dr), nat.ptr(pw), nat.ptr(ext), 0, B,
                 nat.stream_handle(g.device))
        return ext

    def _pinfty(self, device):
        key = str(device)
        cache = getattr(self, "_pinf_cache", None)
        if cache is None or cache[0] != key:
            dirs, w = read_pinfty()
            self._pinf_cache = (key, torch.as_tensor(np.ascontiguousarray(dirs)).to(device),
                                torch.as_tensor(np.ascontiguousarray(w)).to(device))
        return self._pinf_cache[1], self._pinf_cache[2]

    def scattering_batch(self, p, enei_list, X, g, fields=False):
        """Scattering cross sections (B, npol) and differential cross sections
        (B, ndir, npol) over the 256-direction far-field sphere from solutions
        X (B, n, 8P); with ``fields`` also the far fields e, h (B, ndir, 3, P)."""
        enei_list = np.atleast_1d(enei_list)
        B, n, P = len(enei_list), p.n, self.npol
        if P > 8:
            raise ValueError("scattering: at most 8 polarisations per call")
        inm, outm, pol, dr = self._side_flags(p, g)
        pw = self._pw(p, enei_list, g.device)
        dirs, w = self._pinfty(g.device)
        nd = dirs.shape[0]
        dsca = torch.empty((B, nd, P), dtype=torch.float64, device=g.device)
        sca = torch.empty((B, P), dtype=torch.float64, device=g.device)
        E = H = None
        if fields:
            E = torch.empty((B, nd, 3, P), dtype=torch.complex128, device=g.device)
            H = torch.empty_like(E)
        nat.call("mnpbem_scattering_batched", n, P, nat.ptr(X), 8 * P, n * 8 * P, nat.ptr(g.pos), nat.ptr(g.area),
                 nat.ptr(inm), nat.ptr(outm), nd, nat.ptr(dirs), nat.ptr(w), nat.ptr(pw), nat.ptr(dsca),
                 nat.ptr(sca), nat.ptr(E), nat.ptr(H), B, nat.stream_handle(g.device))
        return (sca, dsca, E, H) if fields else (sca, dsca)

    def scattering(self, sig, green=None):
        """Scattering cross section (planewave_ret.py:391-435): (sca, dsca)
        with sca a float for one polarisation, else (npol,); dsca a CompStruct
        holding the differential cross section over the far-field sphere."""
        g = green if green is not None else _green_for(sig.p)
        X = sig.get("_X")
        if X is None:
            raise ValueError("scattering: CompStruct does not come from BEMRet.solve")
        sca, dsca = self.scattering_batch(sig.p, [sig.enei], X, g)
        s = sca[0].cpu().numpy()
        d = dsca[0].cpu().numpy()
        if self.npol == 1:
            return float(s[0]), CompStruct(sig.p, sig.enei, dsca=d[:, 0])
        return s, CompStruct(sig.p, sig.enei, dsca=d)

    def absorption(self, sig, green=None):
        """Absorption cross section ext - sca (planewave_ret.py:363-389)."""
        ext = self.extinction(sig, green)
        sca, _ = self.scattering(sig, green)
        return ext - sca

    def extinction(self, sig, green=None):
        """Optical-theorem extinction cross section (planewave_ret.py:437-504):
        float for one polarisation, else an array (npol,)."""
        g = green if green is not None else _green_for(sig.p)
        X = sig.get("_X")
        if X is None:
            raise ValueError("extinction: CompStruct does not come from BEMRet.solve")
        ext = self.extinction_batch(sig.p, [sig.enei], X, g)[0].cpu().numpy()
        return float(ext[0]) if self.npol == 1 else ext


def _green_for(p):
    from .greenfun import green_for
    return green_for(p)


def read_pinfty(path=None):
    """Directions and weights of the far-field sphere (the pinfty256.bin file
    of the reference: int32 count n, then the n unit vectors as three float64
    arrays x[n], y[n], z[n], then either positions (3n) and weights (n) or the
    n weights alone)."""
    import os
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "pinfty256.bin")
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    vals = raw[4:].view(np.float64)
    dirs = np.stack([vals[0:n], vals[n:2 * n], vals[2 * n:3 * n]], axis=1)
    rest = vals[3 * n:]
    w = rest[3 * n:4 * n] if rest.size == 4 * n else rest[:n]
    return np.ascontiguousarray(dirs), np.ascontiguousarray(w)
