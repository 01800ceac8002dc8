"""Fatigue slab loop on the device: slab solve, damage over the slab, element deletion.

The reference's mechanical slab loop (main3d.f90:609-937, without the thermal side) does,
per slab:

1. element status from the damage so far; failed elements are deleted, orphan nodes
   fixed, the matrices and the preconditioner rebuilt (main3d.f90:625-691);
2. the right-hand side: jump term (ktn1 (x) K + mtn1 (x) M) x_prev plus the load
   fextc (x) rext, fixed DOFs zeroed (main3d.f90:803-823);
3. the slab solve (main3d.f90:831-832);
4. the damage update over the slab's nip temporal points from the new solution
   (main3d.f90:880-898: kokkos_damage, kokkos_damage.f90:36-107).

Here every step runs on the device for one load case of one specimen (the damage
state belongs to the specimen): the solve is ``xt_slab_solve``
(FP32-IR or FP64), the solution goes to the damage kernel in the reference's
temporal-major layout (``xt_from_interleaved``), the element status is the device
reduction, and a deletion event edits the matrices on the host (``ElementDeletion``)
and puts their values into the device operator and banded factor (refactored on the
device from the first changed row on, xt_prec_refresh_band).  The damage law's
constants are parameters (the reference's by default, damage.h:9-17).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import api, device, fem, slab
from ._lib import check, lib
from .deletion import ElementDeletion


class FatigueSlabs:
    """One specimen's slab sequence with damage and deletion.

    mesh, K, M: the assembled matrices before boundary conditions (1-based CSR);
    Nxyz: the shape-derivative tables [8*ne, 8]; bc: per-DOF clamp mask; rext: the
    spatial load vector [p] (traction for amplitude 1); amp, freq: load amplitude and
    frequency; dt: slab length (a whole number of load periods); nip: temporal points of
    the damage update per slab; Dc: critical damage; ym, nu: the material; params: the
    damage law's constants (sig_f, sig_u, eps_D, Cy, h, S, s) or None; mode: solver
    mode of xt_slab_solve (1 FP32-IR, 0 FP64); readd: keep the reference's re-assembly
    of the deleted elements' stiffness (True) or leave them removed (False)."""

    def __init__(self, mesh, K, M, Nxyz, bc, rext, amp, freq, dt, nip, Dc, ym, nu,
                 params=None, mode=1, readd=False, xtol=0.0, im=30, eps=1e-8, refresh=True):
        """refresh: after a deletion event refactor the banded factor from the first
        changed row on (xt_prec_refresh_band); False: refactor it whole."""
        self.mesh, self.nip, self.Dc = mesh, int(nip), float(Dc)
        self.refresh, self.first_row = bool(refresh), 0
        self.mode, self.xtol, self.im, self.eps = mode, xtol, im, eps
        self.p = 3 * mesh.nn
        self.amp, self.w, self.dt = float(amp), 2 * np.pi * float(freq), float(dt)
        c = slab.stcoeff_m(0.0, self.dt, self.w)        # t1 = 0 as main3d.f90:816 passes it
        self.A, self.B = c["k"] + c["ktn"], c["m"] + c["mtn"]
        self.J = np.stack([c["ktn1"], c["mtn1"]])[None]
        self.deletion = ElementDeletion(mesh, K, M, bc, ym, nu, Dc=Dc, readd=readd)
        self.damage = device.DamageState(mesh, *Nxyz, ym, nu, params=params)
        dev = torch.device("cuda")
        self.rext = torch.as_tensor(np.ascontiguousarray(rext, dtype=np.float64), device=dev)
        self.b = torch.zeros(6 * self.p, dtype=torch.float64, device=dev)
        self.x = torch.zeros_like(self.b)
        self.xprev = torch.zeros_like(self.b)
        self.sol = torch.zeros_like(self.b)          # temporal-major copy for the damage kernel
        self._build(self.deletion.K, self.deletion.M, self.deletion.fixed)
        self.n = 0
        self.log = []
        self.separated = False

    def _build(self, K, M, fixed):
        """Device operator, banded factor and fixed mask of the current matrices, stored in
        the element-coupling patterns of K and M (every coupling the elements make, stored
        zeros included: the assembly and the 1-0 method drop entries that sum to zero,
        which a deletion can make nonzero): every matrix a deletion event leaves fits them,
        so after the first build the new values go into the existing handles
        (xt_op_set_values, and xt_prec_refresh_band from the device CSR: the factor from
        the first changed row on) with no host-side operator build or upload."""
        self.fixed = torch.as_tensor((np.asarray(fixed) != 0).astype(np.uint8), device="cuda")
        if getattr(self, "op", None) is None:
            self.patK = fem.dof_pattern(self.mesh, "all")
            self.patM = fem.dof_pattern(self.mesh, "equal")
            kv = api.csr_values_in_pattern(self.patK, K)
            mv = api.csr_values_in_pattern(self.patM, M)
            self.op = device.KronOperator([self.patK + (kv,), self.patM + (mv,)], q=6)
            self.pr = device.BandPreconditioner(self.patK + (kv,))
            self.patK_dev = tuple(torch.as_tensor(v, device="cuda") for v in self.patK)
            self.tj = self.op.tmat(self.J, False)
            self.kv = kv
            return
        kv = api.csr_values_in_pattern(self.patK, K)
        kd = torch.as_tensor(kv, device="cuda")
        md = torch.as_tensor(api.csr_values_in_pattern(self.patM, M), device="cuda")
        self.op.set_values(kd, md)
        # the deletion edits (element matrices removed, orphan diagonals, 1-0 method) are
        # symmetric, so K is unchanged in the rows and columns before its first changed
        # row: the factor is refreshed from there on
        self.first_row = self.refresh_row(self.patK[0], self.kv, kv)
        if self.first_row < self.p:
            self.pr.refactor(self.patK_dev[0], self.patK_dev[1], kd,
                             first_row=self.first_row if self.refresh else 0)
        self.kv = kv

    @staticmethod
    def refresh_row(rowptr, old, new):
        """First (0-based) row of a 1-based CSR whose values differ between old and new
        (len(rowptr) - 1 when none does)."""
        d = np.flatnonzero(np.asarray(old) != np.asarray(new))
        if d.size == 0:
            return len(rowptr) - 1
        return int(np.searchsorted(np.asarray(rowptr), d[0] + 1, side="right") - 1)

    def step(self):
        """One slab: deletion check, right-hand side, solve, damage.  Returns the slab's
        record (iterations, relative residual, newly failed elements, timings); when the
        deletion leaves K singular (the specimen is cut through) the record says
        ``separated`` and no solve runs (nor may later steps)."""
        if self.separated:
            raise RuntimeError("the specimen has separated: no further slabs")
        rec = {"slab": self.n}
        t0 = time.perf_counter()
        ev = self.deletion.step(self.damage)           # chkelesta + edits (main3d.f90:625-687)
        rec["failed"] = [] if ev is None else ev.nefail.tolist()
        self.first_row = 0
        if ev is not None:
            try:
                self._build(ev.K, ev.M, ev.fixed)
            except api.XtError as e:
                if "code -3" not in str(e):
                    raise
                # K singular: the deleted elements cut the specimen in two (a part is no
                # longer held by the clamp); the specimen has failed, the loop ends here
                self.separated = True
                rec["separated"] = True
                self.log.append(rec)
                return rec
        torch.cuda.synchronize()
        rec["deletion_ms"] = 1e3 * (time.perf_counter() - t0)
        if ev is not None:
            rec["refresh_row"] = self.first_row
        t0 = time.perf_counter()
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.op.apply(self.tj, self.xprev, self.b, 1)  # jump term (main3d.f90:818-819)
        fx, _ = slab.getfextc_m(self.n * self.dt, self.w, self.dt)
        coef = torch.as_tensor(np.asarray(fx, dtype=np.float64) * self.amp, device="cuda")
        check(lib().xt_slab_rhs(self.p, 6, 1, coef.data_ptr(), self.rext.data_ptr(), 0,
                                self.fixed.data_ptr(), self.b.data_ptr(), s), "slab_rhs")
        self.x.zero_()
        _, its, ierr, relres = device.slab_solve_device(
            self.op, self.pr, self.A[None], self.B[None], self.b, self.x, 1, self.im, self.eps,
            500, self.mode, self.xtol)
        # the solution at the fixed DOFs is zero (main3d.f90:842-846); the damage kernel
        # reads the temporal-major layout
        check(lib().xt_from_interleaved(self.p, 6, 1, self.x.data_ptr(), self.sol.data_ptr(), s),
              "from_interleaved")
        t1 = self.n * self.dt
        Nt = torch.as_tensor(slab.damage_nt(np.array([t1, t1 + self.dt / 2, t1 + self.dt]),
                                            self.w, self.nip).ravel(), device="cuda")
        self.damage.step(self.sol, Nt, self.nip)
        torch.cuda.synchronize()
        rec.update(its=int(its[0]), ierr=int(ierr[0]), relres=float(relres[0]),
                   solve_damage_ms=1e3 * (time.perf_counter() - t0))
        self.xprev, self.x = self.x, self.xprev
        self.n += 1
        self.log.append(rec)
        return rec

    def max_damage(self):
        return float(self.damage.s["D"].max())

    def deleted(self):
        """All deleted elements so far (1-based, in deletion order)."""
        return np.concatenate([e.nefail for e in self.deletion.events]) if self.deletion.events \
            else np.zeros(0, dtype=np.int32)
