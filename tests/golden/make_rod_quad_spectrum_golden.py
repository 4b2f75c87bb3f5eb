"""Write golden extinction / scattering values of the reference's own
BEMRet on its quadrilateral rod validation case, computed by the reference
code as it stands (mnpbem/bem/bem_ret.py with trirod(10, 40, n=[15, 15, 15]),
ComParticle(..., interp='curv'), tabulated gold; plane waves x-pol along z and
z-pol along x; the set-up of validation/03_bemret/rod/run_python.py) at a few
wavelengths.  The committed validation CSV of that case was written by an
earlier state of the reference: its values differ from what the present code
computes (and agree with MATLAB's file), so this fixture records the present
code's values.

Run once in a container that holds the reference tree (it imports the
reference's numpy code; nothing of the reference is copied):

    python tests/golden/make_rod_quad_spectrum_golden.py /root/reference

Output: tests/golden/rod_quad_spectrum.npz with enei, ext (nlam, 2), sca (nlam, 2).
"""

import os
import sys

import numpy as np

ENEI = [400.0, 600.0, 800.0]


def main(ref_root):
    sys.path.insert(0, ref_root)
    from mnpbem.bem import BEMRet
    from mnpbem.geometry import ComParticle, trirod
    from mnpbem.materials import EpsConst, EpsTable
    from mnpbem.simulation import PlaneWaveRet
    rod = trirod(10.0, 40.0, n=[15, 15, 15])
    p = ComParticle([EpsConst(1.0), EpsTable("gold.dat")], [rod], [[2, 1]], 1, interp="curv")
    bem = BEMRet(p)
    exc = PlaneWaveRet(np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]), np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0]]))
    ext, sca = [], []
    for lam in ENEI:
        sig, bem = bem.solve(exc.potential(p, lam))
        ext.append(np.real(np.ravel(exc.extinction(sig))))
        s = exc.scattering(sig)
        sca.append(np.real(np.ravel(s[0] if isinstance(s, tuple) else s)))
        print(lam, ext[-1], sca[-1])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rod_quad_spectrum.npz")
    np.savez_compressed(path, enei=np.array(ENEI), ext=np.array(ext), sca=np.array(sca))
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
