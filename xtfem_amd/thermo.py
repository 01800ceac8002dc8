"""Thermo-mechanical slab loop on the device (main3d.f90:579-840, damage and element
deletion left to their own modules).

Each slab solves the heat equation on a 3-node temporal slab, updates E and nu at every
Gauss point from the temperature at the slab's end (material_update,
thermoelastic.f90:1-46), re-assembles the mechanical K and M with the thermal load vv
(mpiglbmtx_m, mpisubs.f90:69-139), refactors the preconditioner and solves the
mechanical slab (6 temporal functions) with vv on the right-hand side.  Everything
stays in HBM between slabs: the solutions, the assembly buffers, both operators and
both factors.  Per slab the host only forms the 3x3 and 6x6 temporal data.

Setup (main3d.f90:365-453, 579-608): K_t, M_t with the 1-0 method on the fixed
temperatures and without it; fInit = (A_t (x) K_t0 + B_t (x) M_t0) T_ref and
fBound = -(A_t (x) K_t0 + B_t (x) M_t0) T_fixed with A_t = k + ktn, B_t = m + mtn.
Both temporal slab matrices are taken at t1 = 0: the driver passes the integer slab
index to stcoeff's real t1 (main3d.f90:717, 817), which reads as a subnormal, and
stcoeff_t ignores t1.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import check, lib
from . import api, fem, slab
from .device import (Assembly, BandPreconditioner, KronOperator, _dev, _stream,
                     slab_solve_device)

# Driver material of the parity configuration (main3d.f90:145-157, 193): default-real
# literals widened to FP64 where the reference writes them so.
YM0 = 1.97e5
NU0 = 0.3
RHO = float(np.float32(7.860e-6))
KAPPA = 0.1
CP = 452.0


def driver_conditions(mesh, le=130.0, half_width=35.0, p0=70.0, area_m=664.0, q0=2.0,
                      area_t=13000.0):
    """The boundary conditions and loads main3d.f90 sets up for its panel meshes:
    displacements fixed on z = 0 (and z = 10 le, main3d.f90:482-488), temperatures fixed
    on z = 0 and z = le (429-440), the traction p0 area_m in z lumped per element-node
    incidence on z = le (504-526), the heat flux q0 area_t counted on y = +half_width and
    applied on y = +-half_width (454-478).  Returns (fixed_m [3nn], tfixed [nn],
    rext_m [3nn], rext_t [nn])."""
    fixed_m = fem.clamp_mask(mesh, 2, [0.0, 10.0 * le], tol=0.0)
    tfixed = fem.fixed_temperature_mask(mesh, 2, [0.0, le])
    rext_m = fem.traction_load(mesh, 2, le, 2, p0 * area_m, tol=0.0)
    rext_t = fem.heat_flux_load(mesh, 1, half_width, [half_width, -half_width], q0 * area_t)
    return fixed_m, tfixed, rext_m, rext_t


class ThermoMechanicalSlabs:
    """Device state of the coupled loop for one mesh.

    mesh: fem.Mesh; fixed_m [3*nn] and tfixed [nn]: fixed displacement / temperature
    masks; rext_m [3*nn], rext_t [nn]: the lumped traction and heat-flux vectors; dt:
    slab length; w: angular frequency of the traction (angvel_m).  ``mode`` selects the
    slab solver (0 FP64 GMRES, 1 FP32-inner GMRES iterative refinement).

    The thermal slab is preconditioned by A_t^-1 (x) (K_t + s M_t)^-1 with the exact
    band factor.  ``thermal_shift`` = 0 is the reference's K_t^-1 (mumps_solver on K_t,
    main3d.f90:549); "auto" takes s = the real eigenvalue of A_t^-1 B_t, which accounts
    for the capacity term and halves the GMRES iterations at slabs of 0.5 s.  The
    converged solution is the same to the solver tolerance either way.  ``thermal_block``
    sets the row-block size of that factor (at least the bandwidth); the thermal matrices do
    not change between slabs, so larger blocks (fewer, larger sweep steps) cost only
    setup time."""

    def __init__(self, mesh, fixed_m, tfixed, rext_m, rext_t, dt, w, *, ym0=YM0, nu0=NU0,
                 rho=RHO, kappa=KAPPA, cp=CP, alpha=fem.ALPHA_REF, Tref=0.0, T0=0.0,
                 im_t=100, eps_t=1e-10, maxits_t=500, im_m=100, eps_m=1e-10, maxits_m=500,
                 mode=0, thermal_shift="auto", thermal_block=768, profile=False):
        dev = _dev()
        self.mesh, self.nn, self.p = mesh, mesh.nn, 3 * mesh.nn
        self.dt, self.w = float(dt), float(w)
        self.ym0, self.nu0, self.rho = float(ym0), float(nu0), float(rho)
        self.alpha, self.Tref, self.T0 = float(alpha), float(Tref), float(T0)
        self.solver_t = (min(int(im_t), 256), float(eps_t), int(maxits_t))
        self.solver_m = (min(int(im_m), 256), float(eps_m), int(maxits_m))
        self.mode = int(mode)
        fixed_m = np.asarray(fixed_m) != 0
        tfixed = np.asarray(tfixed) != 0
        if fixed_m.shape != (self.p,) or tfixed.shape != (self.nn,):
            raise api.XtError("fixed_m needs 3*nn entries and tfixed nn entries")
        t64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        u8 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.uint8), device=dev)
        self.fixed_m, self.tfixed = u8(fixed_m), u8(tfixed)
        self.rext_m, self.rext_t = t64(rext_m), t64(rext_t)
        # spatial matrices
        self.asm = Assembly(mesh, fixed_m.astype(np.int32))
        Kt, Mt = self.asm.thermal(tfixed.astype(np.int32), kappa, cp, rho)
        tp = self.asm.T_pattern
        host = lambda v: (tp[0], tp[1], v.cpu().numpy())
        self.op_t = KronOperator([host(Kt), host(Mt)], q=3)
        ct = slab.stcoeff_t(0.0, self.dt)
        At, Bt = ct["k"] + ct["ktn"], ct["m"] + ct["mtn"]
        if thermal_shift == "auto":
            ev = np.linalg.eigvals(np.linalg.solve(At, Bt))
            real = ev.real[np.abs(ev.imag) <= 1e-12 * np.abs(ev).max()]
            shift = float(real.max()) if len(real) else float(ev.real.mean())
        else:
            shift = float(thermal_shift)
        self.thermal_shift = shift
        kv, mv = Kt.cpu().numpy(), Mt.cpu().numpy()
        self.pr_t = BandPreconditioner((tp[0], tp[1], kv + shift * mv if shift else kv),
                                       block=thermal_block)
        op_t0 = KronOperator([host(self.asm.Kt0), host(self.asm.Mt0)], q=3)
        K, M = self.asm.run(self.ym0, self.nu0, self.rho)
        Kp = (self.asm.K_pattern[0], self.asm.K_pattern[1], K.cpu().numpy())
        Mp = (self.asm.M_pattern[0], self.asm.M_pattern[1], M.cpu().numpy())
        self.op_m = KronOperator([Kp, Mp], q=6)
        self.pr_m = BandPreconditioner(Kp)
        # temporal data (t1 = 0, see the module docstring)
        cm = slab.stcoeff_m(0.0, self.dt, self.w)
        self.At, self.Bt = ct["k"] + ct["ktn"], ct["m"] + ct["mtn"]
        self.Am, self.Bm = cm["k"] + cm["ktn"], cm["m"] + cm["mtn"]
        self.jump_t = self.op_t.tmat(np.stack([ct["ktn1"], ct["mtn1"]])[None], False)
        self.jump_m = self.op_m.tmat(np.stack([cm["ktn1"], cm["mtn1"]])[None], False)
        # fInit, fBound (main3d.f90:603-608), interleaved [nn][3]
        sys_t = op_t0.tmat(np.stack([self.At, self.Bt])[None], False)
        self.tempst = torch.full((3 * self.nn,), self.Tref, dtype=torch.float64, device=dev)
        self.finit = torch.zeros_like(self.tempst)
        op_t0.apply(sys_t, self.tempst, self.finit, 1)
        self.tfix0 = torch.zeros(self.nn, dtype=torch.float64, device=dev)
        self.tfix0[self.tfixed.bool()] = self.T0
        self.tempst = self.tfix0.repeat_interleave(3)
        self.fbound = torch.zeros_like(self.tempst)
        op_t0.apply(sys_t, self.tempst, self.fbound, 1, alpha=-1.0)
        del op_t0
        self.dispst = torch.zeros(6 * self.p, dtype=torch.float64, device=dev)
        self.bt = torch.zeros_like(self.tempst)
        self.bm = torch.zeros_like(self.dispst)
        self.coef = torch.zeros(12, dtype=torch.float64, device=dev)
        self.tcoorde = np.zeros(3)
        self.slab = 0
        self.tmask = self.tfixed.bool().repeat_interleave(3)
        self.mmask = self.fixed_m.bool().repeat_interleave(6)
        self.last = {}
        self.profile = bool(profile)
        self._marks = []

    PHASES = ("thermal_rhs", "thermal_solve", "thermo_load", "assembly", "operator_update",
              "band_refactor", "mech_rhs", "mech_solve")

    def _mark(self):
        if self.profile:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self._marks.append(e)

    def phase_ms(self):
        """Per-phase device time (ms) of the last ``step`` when built with profile=True,
        from HIP events on the stream the work runs on."""
        if len(self._marks) != len(self.PHASES) + 1:
            return {}
        self._marks[-1].synchronize()
        return {k: a.elapsed_time(b) for k, a, b in zip(self.PHASES, self._marks[:-1],
                                                        self._marks[1:])}

    def _rhs(self, p, q, coef, rext, fixed, b, slot=0):
        c = self.coef[6 * slot:6 * slot + q]
        c.copy_(torch.as_tensor(np.asarray(coef, dtype=np.float64)))
        check(lib().xt_slab_rhs(p, q, 1, c.data_ptr(), rext.data_ptr(), 0,
                                fixed.data_ptr() if fixed is not None else None, b.data_ptr(),
                                _stream()), "slab_rhs")

    def step(self):
        """One slab: thermal solve, material update and re-assembly, mechanical solve.
        Returns (thermal its, mechanical its)."""
        self.slab += 1
        dt = self.dt
        self.tcoorde = self.tcoorde[2] + np.arange(3) * dt / 2
        t1 = float(self.tcoorde[0])
        self._marks = []
        self._mark()
        # thermal right-hand side (main3d.f90:703-723)
        bt = self.bt
        bt.copy_(self.fbound)
        if self.slab == 1:
            bt.add_(self.finit)
        self.op_t.apply(self.jump_t, self.tempst, bt, 1, beta=1.0)
        self._rhs(self.nn, 3, slab.getfextc_t(t1, dt), self.rext_t, self.tfixed, bt)
        self.tempst.zero_()
        self._mark()
        im, eps, maxits = self.solver_t
        _, its_t, ierr_t, rel_t = slab_solve_device(self.op_t, self.pr_t, self.At[None],
                                                    self.Bt[None], bt, self.tempst, 1, im, eps,
                                                    maxits, self.mode)
        if ierr_t[0] != 0:
            raise api.XtError(f"thermal slab {self.slab}: GMRES ierr={int(ierr_t[0])}")
        self.tempst[self.tmask] = self.T0
        self._mark()
        # material, thermal load and re-assembly (main3d.f90:752-783)
        ym, nu, vv = self.asm.thermo_load(self.tempst, 3, 2, self.Tref, self.alpha, self.ym0,
                                          self.nu0)
        self._mark()
        K, M = self.asm.run(ym, nu, self.rho)
        self._mark()
        self.op_m.set_values(K, M)
        self._mark()
        self.pr_m.refactor(self.asm.K_rowptr, self.asm.K_col, K)
        self._mark()
        # mechanical right-hand side (main3d.f90:803-823)
        fe, fi = slab.getfextc_m(t1, self.w, dt)
        bm = self.bm
        self.op_m.apply(self.jump_m, self.dispst, bm, 1, beta=0.0)
        self._rhs(self.p, 6, fe, self.rext_m, None, bm)
        self._rhs(self.p, 6, fi, vv, self.fixed_m, bm, slot=1)
        self.dispst.zero_()
        self._mark()
        im, eps, maxits = self.solver_m
        _, its_m, ierr_m, rel_m = slab_solve_device(self.op_m, self.pr_m, self.Am[None],
                                                    self.Bm[None], bm, self.dispst, 1, im, eps,
                                                    maxits, self.mode)
        if ierr_m[0] != 0:
            raise api.XtError(f"mechanical slab {self.slab}: GMRES ierr={int(ierr_m[0])}")
        self.dispst[self.mmask] = 0.0
        self._mark()
        self.last = dict(relres_t=float(rel_t[0]), relres_m=float(rel_m[0]))
        return int(its_t[0]), int(its_m[0])

    def write_results(self, files, nip, nipc, damage=None):
        """One output step of the last slab into ``files`` (postprocess.ResultFiles),
        as the driver writes it every nout slabs (main3d.f90:913-917): displacement,
        temperature, strain and stress at the output point, and D, ws and the element
        status from ``damage`` (a device.DamageState; zeros and all active without)."""
        ne = self.mesh.ne
        Nx, Ny, Nz = (t.cpu().numpy().reshape(8 * ne, 8) for t in self.asm.Nxyz)
        if damage is not None:
            st = damage.host_state()
            D, ws = st["D"], st["ws"]
            gpsta = damage.gpsta.cpu().numpy()
            elsta = damage.elesta.cpu().numpy()
        else:
            D = ws = np.zeros(8 * ne)
            gpsta, elsta = np.ones(8 * ne, np.int32), np.ones(ne, np.int32)
        return files.write_step(self.mesh, Nx, Ny, Nz, self.asm.ym_t.cpu().numpy(),
                                self.asm.nu_t.cpu().numpy(), self.w, self.tcoorde,
                                self.displacement(), self.temperature(), gpsta, elsta, D, ws,
                                nip, nipc)

    def temperature(self):
        """tempst, temporal-major [3*nn] on the host."""
        return self.tempst.view(self.nn, 3).t().reshape(-1).cpu().numpy()

    def displacement(self):
        """dispst, temporal-major [18*nn] on the host."""
        return self.dispst.view(self.p, 6).t().reshape(-1).cpu().numpy()
