"""Element deletion at the top of a slab (driver block main3d.f90:625-691).

When some Gauss point reaches the critical damage Dc the reference, in order:

1. marks the failed elements and their Gauss points (chkelesta, on the device
   state here: ``DamageState.failed_elements``);
2. subtracts their stiffness from K (delelement, delelement.f90:1-59);
3. counts the surviving elements of each node and collects the nodes left with
   none (chkcnsta / chkndsta, delelement.f90:139-201);
4. puts a unit diagonal on the DOFs of those orphan nodes (adddiagonal,
   delelement.f90:204-247) and fixes them in every temporal block
   (main3d.f90:666-675);
5. applies the fixed set to M and K (csrapplybc) and refactors the
   preconditioner (main3d.f90:683-687);
6. shortens the slab: dt = dt0*n2, nip = nipc*n2 (main3d.f90:689-690).

The sparse edits are host C++ in libxtfem_hip.so (bit-identical to the
reference's CSR); the caller rebuilds its device operator and preconditioner
from the matrices returned (``KronOperator``, ``BandPreconditioner``), which is
where the time goes.

The reference re-assembles K and M from every element later in the same slab
(main3d.f90:752-783: mpiglbmtx_m over all elements, then csrapplybc with the
updated fixed set), so the matrices its slab solve uses hold the deleted
elements again, with the orphan DOFs fixed.  ``readd=True`` (default) follows
that; ``readd=False`` keeps the edited matrices, so deleted elements stay
removed.  With ``readd=True`` the edited K (which the reference factors and then
replaces) is computed only when ``edited=True``.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import api, fem


@dataclass
class DeletionEvent:
    nefail: np.ndarray            # failed element ids (1-based, increasing)
    ndfail: np.ndarray            # orphan node ids (1-based, discovery order)
    K_edited: tuple               # K after delelement/adddiagonal/csrapplybc (main3d.f90:682),
    M_edited: tuple               # M after csrapplybc (main3d.f90:681); None if not computed
    K: tuple                      # matrices the slab's solve uses (see ``readd``)
    M: tuple
    fixed: np.ndarray             # per-DOF fixed mask after the orphan constraints
    timings: dict = field(default_factory=dict)


class ElementDeletion:
    """Deletion state of one mesh: node survivor counts, fixed DOFs, the current K and M.

    K, M: the assembled matrices before boundary conditions (1-based CSR), as
    ``fem.assemble`` returns them; fixed: per-DOF mask of the clamped DOFs.
    ym, nu: the material; delelement uses the first Gauss point's 6x6 matrix for
    every failed element (the driver passes stiff(:,:,1), main3d.f90:650-652).
    """

    def __init__(self, mesh, K, M, fixed, ym, nu, Dc=0.3, readd=True, edited=False):
        self.mesh = mesh
        self.Dc = float(Dc)
        self.readd = bool(readd)
        self.edited = bool(edited) or not self.readd
        self.K_asm, self.M_asm = K, M
        self.fixed = (np.asarray(fixed) != 0).astype(np.int32)
        self.K = api.csrapplybc(K, self.fixed)
        self.M = api.csrapplybc(M, self.fixed)
        self.necnt = fem.node_degree(mesh).astype(np.int32)     # main3d.f90:273-280
        self.wqs, self.dns, self.dnxyz = fem.shape_tables_m()
        ym1 = float(np.ravel(ym)[0])
        nu1 = float(np.ravel(nu)[0])
        self.stiff = fem.elastic_dmat(ym1, nu1)
        self.events = []

    def apply(self, nefail):
        """Delete the elements ``nefail`` (1-based ids): main3d.f90:650-687.
        Returns a DeletionEvent, or None when the list is empty."""
        nefail = np.asarray(nefail, dtype=np.int32)
        if nefail.size == 0:
            return None
        t = {}
        xconn = self.mesh.xconn
        t0 = time.perf_counter()
        self.necnt = api.chkcnsta(self.mesh.nn, xconn, nefail, self.necnt)
        ndfail = api.chkndsta(self.mesh.nn, xconn, nefail, self.necnt)
        for node in ndfail:
            self.fixed[3 * (node - 1):3 * node] = 1
        t["node_status"] = time.perf_counter() - t0
        # the 1-0 method is idempotent: without a new orphan node the fixed set is the one
        # the current M (and, re-assembling, K) already carry, so they are their own result
        new_fixed = len(ndfail) > 0
        K_ed = M_ed = None
        if self.edited:
            t0 = time.perf_counter()
            K1 = api.delelement(xconn, self.mesh.xcoord, self.dns, self.dnxyz, fem.SD,
                                self.stiff, self.wqs, nefail, self.K)
            K1 = api.adddiagonal(3, ndfail, K1)
            # delelement changes only the couplings among the failed elements' DOFs; when
            # none of them is fixed and the fixed set is unchanged, K1's fixed rows and
            # columns are the current K's, already in 1-0 form, and K1 holds no exact
            # zero: the 1-0 method would return K1 as it is
            nodes = xconn[nefail - 1].ravel() - 1
            dofs = (3 * nodes[:, None] + np.arange(3)).ravel()
            touches = bool(self.fixed[dofs].any())
            K_ed = api.csrapplybc(K1, self.fixed) if (new_fixed or touches) else K1
            M_ed = api.csrapplybc(self.M, self.fixed) if new_fixed else self.M
            t["edit"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        if self.readd:
            if new_fixed:
                self.K = api.csrapplybc(self.K_asm, self.fixed)
                self.M = api.csrapplybc(self.M_asm, self.fixed)
        else:
            self.K, self.M = K_ed, M_ed
        t["next_matrices"] = time.perf_counter() - t0
        ev = DeletionEvent(nefail, ndfail, K_ed, M_ed, self.K, self.M, self.fixed.copy(), t)
        self.events.append(ev)
        return ev

    def step(self, damage):
        """chkelesta on a device ``DamageState`` and, if any element failed, ``apply``."""
        return self.apply(damage.failed_elements(self.Dc))


def shortened_slab(dt0, nipc, n2):
    """Slab length and temporal points per slab after a deletion (main3d.f90:689-690)."""
    return dt0 * n2, nipc * n2
