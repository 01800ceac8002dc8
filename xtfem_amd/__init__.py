"""MI355X-native space-time solve path of XTFEM (extended space-time FEM).

Hot path (HIP, gfx950, libxtfem_hip.so; C ABI in include/xtfem.h):
  * Kronecker space-time operator  Y = sum_k S_k X T_k^T   (kronecker.f90:19-127)
  * Kronecker ILUT preconditioner  (A (x) LU)^-1           (kronecker.f90:130-178, ilut.f)
  * batched restarted GMRES, FP64 or FP32 GMRES-IR         (kronecker.f90:181-365)
  * two-scale fatigue damage per Gauss point + element status
                                       (kokkos_kernel.h:44-344, delelement.f90:109-136)
Host side: ``api`` mirrors the reference subroutines, ``device`` keeps operators and
state resident in HBM, ``fem``/``slab`` build the inputs of a slab.
"""
from . import fem, slab  # noqa: F401
from ._lib import EXPORTS, XtError, lib  # noqa: F401

__all__ = ["fem", "slab", "lib", "EXPORTS", "XtError"]
