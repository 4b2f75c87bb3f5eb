"""MI355X-native retarded BEM (MNPBEM's BEMRet path) on HIP kernels.

Public API mirrors mnpbem: materials (EpsConst, EpsDrude, EpsTable),
geometry (trisphere, Particle, ComParticle), CompGreenRet, BEMRet,
BEMRetIter (GMRES, optionally over ACA H-matrices: ClusterTree, HMatrix,
ACACompGreenRet), PlaneWaveRet and CompStruct.  Device work runs in libmnpbem_amd.so
(built by ``python -m mnpbem_amd.build``); there is no CPU fallback.
"""

from .materials import EV2NM, EpsConst, EpsDrude, EpsTable
from .geometry import ComParticle, Particle, triellipsoid, trirod, trisphere, trisphere_fibonacci

__all__ = ["EV2NM", "EpsConst", "EpsDrude", "EpsTable", "ComParticle", "Particle", "trisphere",
           "trisphere_fibonacci", "triellipsoid", "trirod",
           "CompGreenRet", "BEMRet", "CompStruct", "SweepGraph", "BEMRetBatch", "BatchSweepGraph", "PlaneWaveRet", "BEMIter", "BEMRetIter", "ClusterTree",
           "HMatrix", "ACACompGreenRet", "make_kaware_fadmiss", "DrudeExtinction",
           "extinction_with_drude_gradient", "ShapeExtinction", "extinction_with_shape_gradient"]


def __getattr__(name):
    # device-side classes are imported lazily so that host-only use (mesh,
    # materials) does not need torch or the HIP library
    if name == "CompGreenRet":
        from .greenfun import CompGreenRet
        return CompGreenRet
    if name in ("BEMRet", "CompStruct", "SweepGraph"):
        from . import bem
        return getattr(bem, name)
    if name in ("BEMRetBatch", "BatchSweepGraph"):
        from . import batch
        return getattr(batch, name)
    if name in ("BEMIter", "BEMRetIter"):
        from . import bem_iter
        return getattr(bem_iter, name)
    if name in ("ClusterTree", "HMatrix", "ACACompGreenRet", "make_kaware_fadmiss"):
        from . import hmatrix
        return getattr(hmatrix, name)
    if name in ("DrudeExtinction", "extinction_with_drude_gradient", "ShapeExtinction",
                "extinction_with_shape_gradient"):
        from . import adjoint
        return getattr(adjoint, name)
    if name == "PlaneWaveRet":
        from .planewave import PlaneWaveRet
        return PlaneWaveRet
    raise AttributeError(name)
