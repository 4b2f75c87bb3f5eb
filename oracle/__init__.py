"""CPU FP64 oracle for the retarded-BEM path (test infrastructure only).

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``mnpbem_amd`` must not import anything from
here; it fails loudly when its HIP library is missing instead of falling back.

Contents
--------
``bem_oracle``  numpy/scipy restatement of MNPBEM's retarded BEM direct solve
                (mesh -> curved geometry -> quadrature -> refinement ->
                Green functions -> BEMRet init/solve -> plane-wave extinction),
                every function citing the reference file:line it follows.
"""

from .bem_oracle import *  # noqa: F401,F403
