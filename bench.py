#!/usr/bin/env python3
"""Space-time slab solves on one MI355X.

Headline workload (the stiffened-panel configuration): the panel 240 x 200 x 2 mm as a
C3D8 grid of 240 x 200 x 2 elements (241 x 201 x 3 = 145,323 nodes, p = 435,969),
clamped on three edges (x = 0, x = 240, y = 0), in-plane tensile traction along y on
the free edge y = 200, steel (E = 2e5 MPa, nu = 0.3, rho = 7.85e-9 t/mm^3), n_t = 6.
32 load cases solved together: traction amplitudes U[20, 120] MPa and load
frequencies U[10, 100] Hz from a fixed seed, each case's slab spanning the whole number
of load periods nearest 0.5 s (dt_r = round(f_r / 2) / f_r), the regime in which the
reference's t1 = 0 temporal matrices (main3d.f90:816) describe every slab.  Every case
has its own temporal matrices A_r, B_r.

One step = one space-time slab for every load case, all on the device:
  jump term   b = (ktn1_r (x) K + mtn1_r (x) M) x_prev_r          (kronspgemv)
  load        b += fextc_r (x) amp_r rext, fixed DOFs zeroed       (main3d.f90:803-823)
  solve       (A_r (x) K + B_r (x) M) x_r = b_r, x0 = 0, GMRES(30) to 1e-8 relative
              residual.
value: slab solves per second of the batched FP32-IR solve (FP32 operator and factor
sweeps inside FP64 GMRES iterative refinement, every case also held to a forward-error
test: last correction <= 1e-7 of the solution), preconditioned by A_r^-1 (x) K^-1 with
each case's dominant temporal mode split onto a shifted factor (K + s M)^-1 (exact
banded factors).  Beside it, on the same slabs: plain FP64 GPU GMRES with the
reference's preconditioner A_r^-1 (x) K^-1 on the whole batch (precision ratio) and
one load case per solve (batching ratio); the displacement of each against an FP64
solve to 1e-12; the Kron-SpMM; the CPU baseline (the oracle's FP64 kronpgmres, one
process per host core).  Setup (assembly, upload, factorizations) is outside the timed
region and reported separately.

--workload plate runs the single-edge-notched plate instead (100 x 46 x 4 C3D8, notch
20 x 1 at mid-height, clamped at y = 0, traction on y = 46, p = 71,205), with
--cases-kind plate (all cases at 20 Hz, traction profiles) or sweep (10-100 Hz).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_MFMA_TFLOPS = {"f32": 157.3, "f64": 78.6}   # dense matrix-core peaks
# vector ALU: FP64 FMA, and FP32 FMA without the packed instructions (disabled in this
# build), one per lane per clock: 78.6 TFLOP/s either way
PEAK_VALU_TFLOPS = 78.6
SEED = 20261202


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cases", type=int, default=32, help="load cases per batched solve")
    ap.add_argument("--workload", choices=("panel", "plate", "multimesh"), default="panel",
                    help="multimesh: only the uneven multi-mesh batch leg")
    ap.add_argument("--grid", type=int, nargs=3, default=None,
                    help="elements per axis (panel: 240 200 2, plate: 100 46 4)")
    ap.add_argument("--notch", type=int, default=20, help="plate: notch length in elements")
    ap.add_argument("--block", type=int, default=0, help="factor block rows (0: bandwidth)")
    ap.add_argument("--cases-kind", choices=("plate", "sweep"), default="sweep",
                    help="plate workload: all cases at 20 Hz (plate) or 10-100 Hz (sweep)")
    ap.add_argument("--xtol", type=float, default=1e-7,
                    help="forward-error test of the FP32-IR refinement (0: residual only)")
    ap.add_argument("--shift-ratio", type=float, default=16.0,
                    help="spread of the dominant temporal eigenvalues one shifted factor covers "
                         "(0: no shifted factors)")
    ap.add_argument("--band-levels", type=int, default=0,
                    help="XT_TUNE_BAND_LEVELS for the factors (1: no pairing levels; "
                         "0: the built-in choice)")
    ap.add_argument("--legs", default="",
                    help="extra legs, comma-separated: rod, reassembly, thermo, driver_panel, "
                         "deletion, damage, fatigue, multimesh (or 'all')")
    ap.add_argument("--fatigue-blocks", type=int, default=40,
                    help="timed cycle blocks of the fatigue leg")
    ap.add_argument("--damage-nip", type=int, default=2560,
                    help="temporal points per slab of the damage leg")
    ap.add_argument("--rod-only", action="store_true",
                    help="only run the prismatic-rod ILUT leg (K timed slabs) and print it")
    ap.add_argument("--damage-only", action="store_true",
                    help="only the damage-kernel leg (60x60x12 block), one JSON line")
    ap.add_argument("--fatigue-only", action="store_true",
                    help="only the fatigue loop leg (notched 60x60x12 block), one JSON line")
    ap.add_argument("--thermo-only", action="store_true",
                    help="only run the thermo-mechanical plate leg and print it")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only the CPU baseline (no GPU), printed as one JSON line")
    ap.add_argument("--cpu-budget", type=float, default=45.0,
                    help="seconds of solving per CPU worker in the CPU baseline")
    ap.add_argument("--cpu-matrices", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--profile-only", action="store_true",
                    help="only the setup and the timed FP32-IR steps (for rocprofv3)")
    ap.add_argument("--spmm-only", action="store_true",
                    help="only the workload's operator and the Kron-SpMM microbenchmark "
                         "(FP32, 1 and 32 load cases), one JSON line (for rocprofv3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU comparison")
    return ap.parse_args()


def panel_mesh(nx=240, ny=200, nz=2):
    """The stiffened-panel grid (1 mm C3D8 elements, z fastest, then y, then x: node
    half-bandwidth (ny + 1)(nz + 1) + nz + 2) and its clamp: x = 0, x = nx, y = 0."""
    from xtfem_amd import fem
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz))
    x = mesh.xcoord
    clamp = (x[:, 0] == 0.0) | (x[:, 0] == float(nx)) | (x[:, 1] == 0.0)
    return mesh, np.repeat(clamp, 3).astype(np.int32)


def load_cases(R, seed, kind="sweep"):
    """Amplitudes U[20, 120] MPa and frequencies U[10, 100] Hz (kind sweep; case 0 at
    80 MPa for the plate) or all 20 Hz (kind plate), with each case's slab length, the
    whole number of load periods nearest 0.5 s."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(20.0, 120.0, R)
    freq = rng.uniform(10.0, 100.0, R) if kind == "sweep" else np.full(R, 20.0)
    dt = np.round(0.5 * freq) / freq
    return amp, freq, dt


class PanelProblem:
    """Mesh, spatial matrices (assembled on the device) and per-case slab data of the
    stiffened panel."""

    def __init__(self, R, grid=(240, 200, 2), host_assembly=False):
        from xtfem_amd import fem, slab
        nx, ny, nz = grid
        self.grid = grid
        self.mesh, self.bc = panel_mesh(nx, ny, nz)
        self.p = 3 * self.mesh.nn
        if host_assembly:     # CPU-only process (the CPU baseline): numpy restatement
            K, M, *_ = fem.assemble(self.mesh, 2.0e5, 0.3, 7.85e-9)
            self.Kb, self.Mb = fem.applybc(K, self.bc), fem.applybc(M, self.bc)
        else:                 # device assembly with the 1-0 method (xt_asm_run)
            from xtfem_amd import device
            asm = device.Assembly(self.mesh, self.bc)
            asm.run(2.0e5, 0.3, 7.85e-9)
            self.Kb, self.Mb = asm.host()
            del asm
        self.amp, self.freq, self.dt = load_cases(R, SEED + 3)
        self.w = 2 * np.pi * self.freq
        c = slab.stcoeff_m(0.0, self.dt, self.w)     # t1 = 0 as main3d.f90:816 passes it
        self.A, self.B = c["k"] + c["ktn"], c["m"] + c["mtn"]
        self.J1, self.J2 = c["ktn1"], c["mtn1"]
        unit = fem.traction_load(self.mesh, 1, float(ny), 1, float(nx * nz))
        self.rext = np.tile(unit, (R, 1))

    def coef(self, step, cases):
        """fextc_r(j) * amp_r for slab `step` (tcoorde(1) = step * dt_r, main3d.f90:806)."""
        from xtfem_amd import slab
        fx, _ = slab.getfextc_m(step * self.dt[cases], self.w[cases], self.dt[cases])
        return (np.atleast_2d(fx) * self.amp[cases, None]).ravel()

    def describe(self):
        nx, ny, nz = self.grid
        return (f"stiffened panel {nx}x{ny}x{nz} mm, C3D8 {nx}x{ny}x{nz} elements, "
                f"{self.mesh.nn} nodes, p={self.p}, 3 DOF/node, n_t=6, clamped on x=0, "
                f"x={nx}, y=0, in-plane traction on y={ny}; {len(self.amp)} load cases, "
                f"amplitudes U[20,120] MPa, frequencies U[10,100] Hz (seed {SEED + 3}), "
                f"slabs of whole load periods nearest 0.5 s")


class Problem:
    """Mesh, spatial matrices and per-case slab data of the notched plate."""

    def __init__(self, nx, ny, nz, notch, R, kind="plate"):
        from xtfem_amd import fem, slab
        keep = np.ones((nx, ny, nz), bool)
        keep[0:notch, ny // 2, :] = False
        self.mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep)
        K, M, *self.Nxyz = fem.assemble(self.mesh, 2.0e5, 0.3, 7.85e-9)
        self.K, self.M = K, M
        self.bc = fem.clamp_mask(self.mesh, 1, [0.0])
        self.Kb, self.Mb = fem.applybc(K, self.bc), fem.applybc(M, self.bc)
        self.p = 3 * self.mesh.nn
        rng = np.random.default_rng(SEED)
        self.amp = np.concatenate([[80.0], rng.uniform(20.0, 120.0, R - 1)])
        if kind == "sweep":
            self.freq = np.concatenate([[20.0], rng.uniform(10.0, 100.0, R - 1)])
        else:
            self.freq = np.full(R, 20.0)
        self.w = 2 * np.pi * self.freq
        self.dt = np.round(0.5 * self.freq) / self.freq
        self.kind, self.grid, self.notch = kind, (nx, ny, nz), notch
        # t1 = 0 in stcoeff_m, as the reference's call passes it (main3d.f90:816)
        c = slab.stcoeff_m(0.0, self.dt, self.w)
        self.A, self.B = c["k"] + c["ktn"], c["m"] + c["mtn"]
        self.J1, self.J2 = c["ktn1"], c["mtn1"]
        # unit traction on y = ny (total force = area of the loaded face), then each
        # case's profile along x; rows of self.rext are per case [R, p]
        unit = fem.traction_load(self.mesh, 1, float(ny), 1, float(nx * nz))
        xnode = np.repeat(self.mesh.xcoord[:, 0], 3)
        self.rext = np.tile(unit, (R, 1))
        if kind == "plate":
            a = rng.uniform(0.0, 0.5, R)
            m = rng.integers(1, 4, R)
            phi = rng.uniform(0.0, 2 * np.pi, R)
            a[0] = 0.0
            for r in range(R):
                self.rext[r] *= 1.0 + a[r] * np.sin(2 * np.pi * m[r] * xnode / nx + phi[r])

    def coef(self, step, cases):
        """fextc_r(j) * amp_r for slab `step` (tcoorde(1) = step * dt_r, main3d.f90:806)."""
        from xtfem_amd import slab
        fx, _ = slab.getfextc_m(step * self.dt[cases], self.w[cases], self.dt[cases])
        return (np.atleast_2d(fx) * self.amp[cases, None]).ravel()

    def describe(self):
        nx, ny, nz = self.grid
        return (f"notched plate {nx}x{ny}x{nz} C3D8, notch {self.notch}x1, 3 DOF/node, "
                f"p={self.p}, n_t=6, {len(self.amp)} load cases ({self.kind})")


class RodProblem(Problem):
    """The prismatic rod: C3D8 4 x 4 x 12 elements over 10 x 10 x 30 mm (325 nodes),
    clamped at z = 0, axial traction 100 MPa at 50 Hz on z = 30, E = 2e5 MPa,
    nu = 0.3, rho = 7.85e-9 t/mm^3, slabs of 0.5 s (25 periods), one load case."""

    def __init__(self):
        from xtfem_amd import fem, slab
        self.mesh = fem.structured_hex(4, 4, 12, 10.0, 10.0, 30.0)
        K, M, *_ = fem.assemble(self.mesh, 2.0e5, 0.3, 7.85e-9)
        self.bc = fem.clamp_mask(self.mesh, 2, [0.0])
        self.Kb, self.Mb = fem.applybc(K, self.bc), fem.applybc(M, self.bc)
        self.p = 3 * self.mesh.nn
        self.amp = np.array([100.0])
        self.freq = np.array([50.0])
        self.w = 2 * np.pi * self.freq
        self.dt = np.array([0.5])
        c = slab.stcoeff_m(0.0, self.dt, self.w)
        self.A, self.B = c["k"] + c["ktn"], c["m"] + c["mtn"]
        self.J1, self.J2 = c["ktn1"], c["mtn1"]
        self.rext = fem.traction_load(self.mesh, 2, 30.0, 2, 100.0)[None]


def rod_leg(K):
    """The rod with the reference's ILUT path: ILUT(20, 1e-4), GMRES(30) to 1e-8, one
    slab per solve, FP64 and FP32-IR on the GPU."""
    from xtfem_amd import device
    prob = RodProblem()
    op = device.KronOperator([prob.Kb, prob.Mb], q=6)
    pr = device.IlutPreconditioner(prob.Kb, lfil=20, droptol=1e-4)
    out = {"workload": f"rod 4x4x12 C3D8 (p={prob.p}), ILUT(20, 1e-4), GMRES(30), eps 1e-8"}
    for mode, name in ((0, "fp64"), (1, "fp32_ir")):
        run = Runner(prob, op, pr, np.array([0]), mode=mode, nsteps=K + 2)
        run.step()
        t = timed(run, K)
        rr = np.array(run.relres[1:])
        out[name] = dict(slabs_per_s=K / t, ms_per_slab=1e3 * t / K,
                         its_max=int(np.max(run.its[1:])), relres_max=float(rr.max()))
    return out


class Runner:
    """Device state of one slab sequence (R load cases) and its step."""

    def __init__(self, prob, op, pr, cases, mode, nsteps, xtol=0.0):
        import torch
        from xtfem_amd import device
        self.prob, self.op, self.pr, self.mode, self.xtol = prob, op, pr, mode, xtol
        self.cases = np.asarray(cases)
        self.R = len(self.cases)
        p, R = prob.p, self.R
        dev = torch.device("cuda")
        self.tj = op.tmat(np.stack([prob.J1[self.cases], prob.J2[self.cases]], axis=1), False)
        self.rext = torch.as_tensor(np.ascontiguousarray(prob.rext[self.cases]), device=dev)
        self.fixed = torch.as_tensor(prob.bc.astype(np.uint8), device=dev)
        self.coefs = torch.as_tensor(np.stack([prob.coef(n, self.cases) for n in range(nsteps)]),
                                     device=dev)
        self.b = torch.zeros(p * 6 * R, dtype=torch.float64, device=dev)
        self.x = torch.zeros_like(self.b)
        self.xprev = torch.zeros_like(self.b)
        self.A, self.B = prob.A[self.cases], prob.B[self.cases]
        self.device = device
        self.eps = 1e-8
        self.n = 0
        self.its, self.relres, self.ierr = [], [], []
        self.it_prec = []

    def step(self):
        from xtfem_amd._lib import check, lib
        import ctypes as C
        import torch
        s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self.op.apply(self.tj, self.xprev, self.b, self.R, alpha=1.0, beta=0.0)
        check(lib().xt_slab_rhs(self.prob.p, 6, self.R, self.coefs[self.n].data_ptr(),
                                self.rext.data_ptr(), self.prob.p, self.fixed.data_ptr(),
                                self.b.data_ptr(), s), "slab_rhs")
        self.x.zero_()
        _, its, ierr, relres = self.device.slab_solve_device(
            self.op, self.pr, self.A, self.B, self.b, self.x, self.R, im=30, eps=self.eps,
            maxits=500, mode=self.mode, xtol=self.xtol)
        self.its.append(its)
        self.ierr.append(ierr)
        self.relres.append(relres)
        self.it_prec.append(self.device.slab_stats(self.op))
        self.xprev, self.x = self.x, self.xprev
        self.n += 1


def timed(run, K):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        run.step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_profile(run, K, mask):
    """Second pass over K steps with per-launch HIP events on the launching stream."""
    import ctypes as C
    from xtfem_amd._lib import check, lib
    check(lib().xt_prof_enable(mask), "prof_enable")
    timed(run, K)
    out = {}
    for kind, name in ((0, "kron_spmm"), (1, "band_sweep"), (3, "damage"),
                       (4, "band_sweep_f64")):
        if not (mask >> kind) & 1:
            continue
        ms, n, by, fl = C.c_double(), C.c_longlong(), C.c_double(), C.c_double()
        check(lib().xt_prof_read(kind, C.addressof(ms), C.addressof(n), C.addressof(by),
                                 C.addressof(fl)), "prof_read")
        ab, af = C.c_double(), C.c_double()
        check(lib().xt_prof_read_alg(kind, C.addressof(ab), C.addressof(af)), "prof_read_alg")
        if n.value:
            # bytes / flops: algorithmic (the unpaired computation, each operand once);
            # performed_*: what the launches do (the paired recurrences add products)
            out[name] = dict(ms=ms.value, launches=n.value, bytes=ab.value, flops=af.value,
                             performed_bytes=by.value, performed_flops=fl.value,
                             avg_us=1e3 * ms.value / n.value,
                             gbs=ab.value / (ms.value * 1e-3) / 1e9,
                             tflops=af.value / (ms.value * 1e-3) / 1e12)
    check(lib().xt_prof_enable(0), "prof_enable")
    return out


def spmm_microbench(op, p, reps=20, warm=3, R=1):
    """Kron-SpMM alone, R load cases, FP32: median per-launch time over reps after
    warm-ups."""
    import ctypes as C
    import torch
    from xtfem_amd import slab
    from xtfem_amd._lib import check, lib
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * 20.0)
    T = op.tmat(np.tile(np.stack([c["k"] + c["ktn"], c["m"] + c["mtn"]])[None], (R, 1, 1, 1)),
                True)
    g = torch.Generator(device="cuda").manual_seed(SEED)
    x = torch.rand(p * 6 * R, device="cuda", generator=g, dtype=torch.float32) * 2 - 1
    y = torch.empty_like(x)
    for _ in range(warm):
        op.apply(T, x, y, R, fp32=True)
    times = []
    for _ in range(reps):
        check(lib().xt_prof_enable(1), "prof_enable")
        op.apply(T, x, y, R, fp32=True)
        ms, n, by, fl = C.c_double(), C.c_longlong(), C.c_double(), C.c_double()
        check(lib().xt_prof_read(0, C.addressof(ms), C.addressof(n), C.addressof(by),
                                 C.addressof(fl)), "prof_read")
        times.append(ms.value)
    check(lib().xt_prof_enable(0), "prof_enable")
    med = float(np.median(times))
    csr = op.algorithmic_bytes(R, True, fmt="csr")
    # which roof bounds the launch: HBM below the ridge intensity (peak flops / peak
    # bytes), the vector ALU above it; attainable_frac = the bound's time / the time taken
    inten = op.flops(R) / by.value
    t_hbm = by.value / (PEAK_HBM_GBS * 1e9)
    t_valu = op.flops(R) / (PEAK_VALU_TFLOPS * 1e12)
    return dict(R=R, dtype="f32", format=op.format, median_us=1e3 * med, bytes=by.value,
                gbs=by.value / (med * 1e-3) / 1e9,
                frac=by.value / (med * 1e-3) / 1e9 / PEAK_HBM_GBS,
                csr_bytes=csr, csr_gbs=csr / (med * 1e-3) / 1e9,
                tflops=op.flops(R) / (med * 1e-3) / 1e12,
                valu_frac=op.flops(R) / (med * 1e-3) / 1e12 / PEAK_VALU_TFLOPS,
                intensity_flop_per_byte=inten,
                ridge_flop_per_byte=PEAK_VALU_TFLOPS * 1e3 / PEAK_HBM_GBS,
                bound="valu" if t_valu > t_hbm else "hbm",
                attainable_frac=max(t_hbm, t_valu) / (med * 1e-3))


def spmm_legs(op, p):
    """Kron-SpMM alone on the workload's operator: one load case (c = 6, HBM-bound) and
    32 (c = 192); bytes of the stored (node-block) format and of canonical CSR."""
    return {f"R{R}": spmm_microbench(op, p, R=R) for R in (1, 32)}


# Algorithmic FP64 operations of one Gauss point and time point of the damage kernel
# in its elastic branch (the count of the kernel's arithmetic: 36 for the shared
# nodal interpolation, 144 strain, 15 damage factors, 31 trial strain, 24 D_el
# product, 13 deviator, 19 von Mises incl. sqrt); the plastic corrector adds ~235 and
# the damage increment ~30, so this is a lower bound when points yield.
DAMAGE_FLOPS_ELASTIC = 282


_DAMAGE = {}


def damage_problem():
    """The damage leg's block, shape-derivative tables and seeded field (built once)."""
    if not _DAMAGE:
        from xtfem_amd import fem
        mesh = fem.structured_hex(60, 60, 12, 60.0, 60.0, 12.0)
        _, _, Nx, Ny, Nz = fem.assemble(mesh, 1.97e5, 0.3, 7.86e-6)
        rng = np.random.default_rng(SEED + 4)
        xi = mesh.xcoord / np.array([60.0, 60.0, 12.0])
        w = rng.uniform(0.5, 1.0, 6) * np.array([1, 1, 1, 0.6, 0.6, 0.6])
        eps0 = 7.0e-4
        sol = np.zeros((6, mesh.nn, 3))
        for k in range(6):
            sol[k, :, 0] = eps0 * w[k] * 60.0 * xi[:, 0] * (1 + 0.3 * xi[:, 1])
            sol[k, :, 1] = (-0.3 * eps0 * w[k] * 60.0 * xi[:, 1]
                            + 0.2 * eps0 * w[k] * 60.0 * xi[:, 0] ** 2)
            sol[k, :, 2] = -0.3 * eps0 * w[k] * 12.0 * xi[:, 2]
        _DAMAGE.update(mesh=mesh, Nx=Nx, Ny=Ny, Nz=Nz, sol=sol.ravel())
    return _DAMAGE


def damage_leg(nip, slabs=3):
    """Damage kernel on the 60 x 60 x 12 C3D8 block (43,200 elements, 345,600 Gauss
    points), nip temporal points per slab, seeded smooth stretch/bend field (axial
    strain 7e-4, peak von Mises about the 180 MPa threshold)."""
    import torch
    from xtfem_amd import slab
    from xtfem_amd.device import DamageState
    dp = damage_problem()
    mesh = dp["mesh"]
    st = DamageState(mesh, dp["Nx"], dp["Ny"], dp["Nz"], 1.97e5, 0.3)
    sol_d = torch.as_tensor(dp["sol"], device="cuda")
    Nt = torch.as_tensor(slab.damage_nt(np.array([0.0, 0.25, 0.5]), 2 * np.pi * 20.0, nip).ravel(),
                         device="cuda")
    st.step(sol_d, Nt, nip)                       # warm-up slab
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(slabs):
        st.step(sol_d, Nt, nip)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / slabs
    ngp = 8 * mesh.ne
    gps = ngp * nip / t
    D = st.s["D"].cpu().numpy()
    yielded = float(np.mean(st.s["p"].cpu().numpy() > 0))
    return dict(workload=f"C3D8 block 60x60x12, {ngp} Gauss points, nip {nip}", ms_per_slab=1e3 * t,
                gp_steps_per_s=gps, tflops_elastic_count=gps * DAMAGE_FLOPS_ELASTIC / 1e12,
                frac_fp64_peak=gps * DAMAGE_FLOPS_ELASTIC / 78.6e12, yielded_fraction=yielded,
                max_D=float(D.max()))


FATIGUE_PARAMS = [350.0, 577.0, 0.08, 1740.0, 0.2, 0.5, 0.5]   # sig_f = 350 MPa yield


def fatigue_leg(blocks=40, nip=2560, amp=250.0):
    """The fatigue-damage configuration: notched block 60 x 60 x 12 mm (C3D8 60 x 60 x 12
    minus a 15 x 1 x 12 notch at mid-height, 43,020 elements), clamped at y = 0, 20 Hz
    tensile traction on y = 60, damage law with a 350 MPa fatigue limit and critical
    damage 0.95.  One cycle block = one slab of 10 load periods: FP32-IR slab solve, the
    damage kernel over nip temporal points of the slab, element status, and deletion with
    the device rebuild when elements fail (xtfem_amd.fatigue: the operator's values
    replaced, the banded factor refreshed from the first changed row on)."""
    import torch
    from xtfem_amd import fem
    from xtfem_amd.fatigue import FatigueSlabs
    nx, ny, nz = 60, 60, 12
    keep = np.ones((nx, ny, nz), bool)
    keep[0:15, ny // 2, :] = False
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep)
    K, M, Nx, Ny, Nz = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    bc = fem.clamp_mask(mesh, 1, [0.0])
    unit = fem.traction_load(mesh, 1, float(ny), 1, float(nx * nz))
    t0 = time.perf_counter()
    fs = FatigueSlabs(mesh, K, M, (Nx, Ny, Nz), bc, unit, amp, 20.0, 0.5, nip, 0.95, 2.0e5,
                      0.3, params=FATIGUE_PARAMS, mode=1, readd=False)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    fs.step()                                          # warm-up block
    recs = []
    t0 = time.perf_counter()
    for _ in range(blocks):
        recs.append(fs.step())
        if fs.separated:
            break
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    recs = [r for r in recs if "its" in r]
    blocks = len(recs)
    dels = [r for r in recs if r["failed"]]
    plain = [r["solve_damage_ms"] for r in recs if not r["failed"]]
    st = fs.damage.host_state()
    return dict(workload=f"notched block {nx}x{ny}x{nz} C3D8 ({mesh.ne} elements, "
                         f"{8 * mesh.ne} Gauss points, p={3 * mesh.nn}), traction {amp} MPa at "
                         f"20 Hz, fatigue limit 350 MPa, Dc 0.95, nip {nip} per block, FP32-IR",
                blocks=blocks, separated=fs.separated, ms_per_block=1e3 * t / max(blocks, 1),
                ms_solve_and_damage_median=float(np.median(plain)) if plain else None,
                deletion_events=len(dels), deleted_elements=int(len(fs.deleted())),
                deletion_ms=[r["deletion_ms"] for r in dels][:8],
                refresh_from_row_frac=[round(r.get("refresh_row", 0) / (3 * mesh.nn), 3)
                                       for r in dels][:8],
                its_max=int(max(r["its"] for r in recs)),
                relres_max=float(max(r["relres"] for r in recs)),
                all_converged=bool(all(r["ierr"] == 0 for r in recs)),
                max_D=float(st["D"].max()), yielded_fraction=float(np.mean(st["p"] > 0)),
                setup_s=t_setup)


def multimesh_meshes(nmesh=64, target_dof=6.0e6):
    """The uneven multi-mesh batch: nmesh prismatic rods (n x n cross-section, length
    6-14 sides) and plates (length 10-25 thicknesses, 2-6 widths), element counts drawn
    log-uniform in [2k, 400k] (seeded) and scaled by one common factor so that the batch
    holds about target_dof spatial DOFs; each clamped at x = 0, its far end x = L held
    transversely (free to stretch) and loaded by an axial traction.  The length bounds
    and the guided far end keep every stiffness matrix's condition number below about
    1e7 (a cantilever 10^4 sides long reaches 10^16, beyond FP64).  Returns
    [(dims, kind)] sorted by node-slice size (x is the slowest node axis, so the
    bandwidth grows with the (ny+1)(nz+1) slice) and the scale factor."""
    rng = np.random.default_rng(SEED + 5)
    ne = np.exp(rng.uniform(np.log(2e3), np.log(4e5), nmesh))
    kind = rng.integers(0, 2, nmesh)
    slender = np.where(kind == 0, rng.uniform(6.0, 14.0, nmesh), rng.uniform(10.0, 25.0, nmesh))
    aspect = rng.uniform(2.0, 6.0, nmesh)

    def dims(i, f):
        e = ne[i] * f
        if kind[i] == 0:
            n = max(2, int(round((e / slender[i]) ** (1.0 / 3.0))))
            return (max(4, int(round(e / (n * n)))), n, n)
        L = (e * aspect[i] * slender[i]) ** (1.0 / 3.0)
        nz = max(2, int(round(L / slender[i])))
        ny = max(nz, int(round(L / aspect[i])))
        return (max(4, int(round(e / (ny * nz)))), ny, nz)

    def total(f):
        return sum(3 * (d[0] + 1) * (d[1] + 1) * (d[2] + 1) for d in (dims(i, f) for i in range(nmesh)))

    lo, hi = 1e-3, 1.0
    for _ in range(60):
        mid = np.sqrt(lo * hi)
        lo, hi = (mid, hi) if total(mid) < target_dof else (lo, mid)
    f = hi if total(hi) <= 1.02 * target_dof else lo
    out = [(dims(i, f), "rod" if kind[i] == 0 else "plate") for i in range(nmesh)]
    out.sort(key=lambda t: ((t[0][1] + 1) * (t[0][2] + 1), t[0][0]))
    return out, f


def consecutive_groups(sizes, budget):
    """Runs of consecutive indices whose sizes sum to at most budget each (an index whose
    size alone exceeds it forms its own group), in order; every index in exactly one."""
    groups, cur, acc = [], [], 0.0
    for i, sz in enumerate(sizes):
        if cur and acc + sz > budget:
            groups.append(cur)
            cur, acc = [], 0.0
        cur.append(i)
        acc += sz
    if cur:
        groups.append(cur)
    return groups


def multimesh_leg(K=3, R=4, nmesh=64, target_dof=6.0e6, xtol=1e-7, group_gb=48.0):
    """The uneven multi-mesh batch: every mesh's K and M stacked block-diagonally, a
    banded factor per run of meshes of similar bandwidth (FP32, xt_prec_create_band_seg),
    one batched FP32-IR slab solve of nmesh x R instances with
    per-instance convergence (xt_slab_solve_seg), timed over K solves; against the same
    meshes solved one at a time in FP64 on the GPU (each with its own operator and FP64
    banded factor, R load cases per solve), and in FP64 batched by groups: consecutive
    meshes (similar bandwidth) whose FP64 factors' four block sets take at most group_gb
    GB together, one segmented FP64 solve per group (the FP64 factors of all meshes do
    not fit in HBM at once)."""
    import torch
    from xtfem_amd import device, fem, slab
    t0 = time.perf_counter()
    spec, scale = multimesh_meshes(nmesh, target_dof)
    amp, freq, dt = load_cases(R, SEED + 6)
    w = 2 * np.pi * freq
    c = slab.stcoeff_m(0.0, dt, w)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    fx, _ = slab.getfextc_m(0.0, w, dt)
    systems = []
    for (nx, ny, nz), kind in spec:
        mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz))
        bc = fem.clamp_mask(mesh, 0, [0.0])
        far = np.flatnonzero(mesh.xcoord[:, 0] == float(nx))
        bc[3 * far + 1] = 1                    # far end guided: y, z held, x free
        bc[3 * far + 2] = 1
        asm = device.Assembly(mesh, bc)
        asm.run(2.0e5, 0.3, 7.85e-9)
        Kb, Mb = asm.host()
        del asm
        rext = fem.traction_load(mesh, 0, float(nx), 0, float(ny * nz))
        rhs = np.array([slab.slab_rhs(3 * mesh.nn, fx[r] * amp[r], None, rext,
                                      fixed=np.tile(bc, 6)) for r in range(R)])
        systems.append(dict(p=3 * mesh.nn, ne=mesh.ne, Kb=Kb, Mb=Mb, rhs=rhs, kind=kind))
    Kc, seg = device.block_diagonal([s["Kb"] for s in systems])
    Mc, _ = device.block_diagonal([s["Mb"] for s in systems])
    t_mesh = time.perf_counter() - t0
    t0 = time.perf_counter()
    op = device.KronOperator([Kc, Mc], q=6)
    pr = device.BandPreconditioner(Kc, fp32_only=True, seg=seg)
    groups = dict(runs=len(pr.group_blocks), blocks=[int(v) for v in pr.group_blocks],
                  factor_gb=pr.matrix_bytes(True) / 1e9)
    b = torch.cat([device.to_interleaved(s["rhs"], s["p"], 6) for s in systems]).contiguous()
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    device.slab_solve_seg(op, pr, seg, A, B, b, x, R, mode=1, xtol=xtol)   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        x.zero_()
        _, its, ierr, relres = device.slab_solve_seg(op, pr, seg, A, B, b, x, R, mode=1,
                                                     xtol=xtol)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / K
    print(f"# multimesh: batched FP32-IR solve {1e3 * t32:.0f} ms", file=sys.stderr, flush=True)
    x32 = x.clone()
    del op, pr
    torch.cuda.empty_cache()
    # one mesh at a time in FP64 (setup excluded from the timed solves)
    t64, its64, err, x64, fbytes = 0.0, [], [], [], []
    for si, s in enumerate(systems):
        op1 = device.KronOperator([s["Kb"], s["Mb"]], q=6)
        pr1 = device.BandPreconditioner(s["Kb"])
        fbytes.append(pr1.matrix_bytes(False))
        b1 = device.to_interleaved(s["rhs"], s["p"], 6)
        x1 = torch.zeros_like(b1)
        device.slab_solve_device(op1, pr1, A, B, b1, x1, R, mode=0)       # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x1.zero_()
        _, it1, ie1, rr1 = device.slab_solve_device(op1, pr1, A, B, b1, x1, R, mode=0)
        torch.cuda.synchronize()
        t64 += time.perf_counter() - t1
        its64 += it1.tolist()
        xs = x32[seg[si] * 6 * R:seg[si + 1] * 6 * R]
        d = (xs - x1).view(-1, 6 * R)
        ref = x1.view(-1, 6 * R)
        for r in range(R):
            cols = slice(6 * r, 6 * r + 6)
            err.append(float(torch.linalg.norm(d[:, cols]) / torch.linalg.norm(ref[:, cols])))
        x64.append(x1)
        del op1, pr1
        if si % 16 == 15:
            print(f"# multimesh: {si + 1} meshes solved one at a time", file=sys.stderr,
                  flush=True)
    del x32
    torch.cuda.empty_cache()
    # FP64 batched by groups of consecutive meshes within group_gb of factor blocks each
    groups64 = consecutive_groups(fbytes, group_gb * 1e9)
    tg, itsg, relg, convg, errg = 0.0, [], [], True, []
    for gi in groups64:
        Kg, segg = device.block_diagonal([systems[i]["Kb"] for i in gi])
        Mg, _ = device.block_diagonal([systems[i]["Mb"] for i in gi])
        opg = device.KronOperator([Kg, Mg], q=6)
        prg = device.BandPreconditioner(Kg, seg=segg)
        bg = torch.cat([device.to_interleaved(systems[i]["rhs"], systems[i]["p"], 6)
                        for i in gi]).contiguous()
        xg = torch.zeros_like(bg)
        device.slab_solve_seg(opg, prg, segg, A, B, bg, xg, R, mode=0)   # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        xg.zero_()
        _, itg, ieg, rrg = device.slab_solve_seg(opg, prg, segg, A, B, bg, xg, R, mode=0)
        torch.cuda.synchronize()
        tg += time.perf_counter() - t1
        itsg += np.asarray(itg).ravel().tolist()
        relg += np.asarray(rrg).ravel().tolist()
        convg = convg and bool(np.all(np.asarray(ieg) == 0))
        for k, si in enumerate(gi):
            xs = xg[segg[k] * 6 * R:segg[k + 1] * 6 * R]
            d = (xs - x64[si]).view(-1, 6 * R)
            ref = x64[si].view(-1, 6 * R)
            for r in range(R):
                cols = slice(6 * r, 6 * r + 6)
                errg.append(float(torch.linalg.norm(d[:, cols]) /
                                  torch.linalg.norm(ref[:, cols])))
        del opg, prg, xg, bg
        torch.cuda.empty_cache()
        print(f"# multimesh: FP64 group of {len(gi)} meshes solved", file=sys.stderr,
              flush=True)
    n_inst = nmesh * R
    elems = [s["ne"] for s in systems]
    return dict(workload=f"{nmesh} meshes ({sum(1 for s in systems if s['kind'] == 'rod')} "
                         f"rods, the rest plates), {min(elems)}-{max(elems)} C3D8 elements "
                         f"(log-uniform draw in [2k, 400k] scaled by {scale:.3f}), "
                         f"{int(seg[-1])} spatial DOFs in all, {R} load cases each, one "
                         f"batched FP32-IR solve of {n_inst} instances",
                fp32_ir_batched_slabs_per_s=n_inst / t32, ms_per_batched_solve=1e3 * t32,
                fp64_one_at_a_time_slabs_per_s=n_inst / t64,
                speedup=(n_inst / t32) / (n_inst / t64),
                its_min=int(its.min()), its_max=int(its.max()),
                fp64_its_max=int(max(its64)),
                all_converged=bool(np.all(ierr == 0) and np.all(relres <= 1e-8)),
                relres_max=float(relres.max()),
                rel_err_vs_fp64_max=float(max(err)),
                fp64_grouped=dict(groups=len(groups64), group_gb=group_gb,
                                  meshes_per_group=[len(g) for g in groups64],
                                  slabs_per_s=n_inst / tg,
                                  speedup_vs_one_at_a_time=(n_inst / tg) / (n_inst / t64),
                                  its_max=int(max(itsg)),
                                  all_converged=bool(convg and max(relg) <= 1e-8),
                                  relres_max=float(max(relg)),
                                  rel_err_vs_one_at_a_time_max=float(max(errg))),
                banded_factor=groups,
                setup_s=dict(meshes=t_mesh, operator_and_factor=t_setup))


def deletion_leg(prob, R=8):
    """Element deletion on the plate (main3d.f90:625-691): damage planted at the notch
    tip (the element column through the thickness) and at a free corner element (its
    corner node becomes an orphan); device element status, host CSR edits
    (delelement, chkcnsta/chkndsta, adddiagonal, 1-0 method), then the device
    operator and banded factor rebuilt from the matrices the next slab uses
    (the reference's re-assembled ones), and one batched slab solve on them."""
    import torch
    from xtfem_amd import device, fem
    from xtfem_amd.deletion import ElementDeletion
    mesh = prob.mesh
    nx, ny, nz = (int(round(mesh.xcoord[:, k].max())) for k in range(3))
    cen = mesh.xcoord[mesh.xconn - 1].mean(axis=1)
    tip_x = max(cen[(np.abs(cen[:, 1] - (ny // 2 + 0.5)) < 0.01), 0].min(), 0.5)
    tip = np.flatnonzero((np.abs(cen[:, 0] - tip_x) < 0.01) &
                         (np.abs(cen[:, 1] - (ny // 2 + 0.5)) < 0.01)) + 1
    deg = fem.node_degree(mesh)
    ytop = mesh.xcoord[:, 1].max()
    corner = [n for n in np.flatnonzero(deg == 1) if mesh.xcoord[n, 1] == ytop][0] + 1
    e_corner = int(np.flatnonzero((mesh.xconn == corner).any(axis=1))[0]) + 1
    failed = np.unique(np.r_[tip, e_corner]).astype(np.int32)
    ds = device.DamageState(mesh, *prob.Nxyz, 2.0e5, 0.3)
    D = np.zeros(8 * mesh.ne)
    D[8 * (failed - 1)] = 0.31
    ds.s["D"].copy_(torch.as_tensor(D, device="cuda"))
    dl = ElementDeletion(mesh, prob.K, prob.M, prob.bc, 2.0e5, 0.3, Dc=0.3, readd=True)
    torch.cuda.synchronize()
    t = {}
    t0 = time.perf_counter()
    nef = ds.failed_elements(0.3)
    t["element_status"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ev = dl.apply(nef)
    t["host_edits"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    op = device.KronOperator([ev.K, ev.M], q=6)
    torch.cuda.synchronize()
    t["operator_upload"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    pr = device.BandPreconditioner(ev.K)
    torch.cuda.synchronize()
    t["band_refactor"] = time.perf_counter() - t0
    cases = np.arange(R)
    saved = prob.bc
    prob.bc = ev.fixed
    run = Runner(prob, op, pr, cases, mode=1, nsteps=2)
    prob.bc = saved
    t0 = time.perf_counter()
    run.step()
    torch.cuda.synchronize()
    t["first_slab_solve"] = time.perf_counter() - t0
    ok = bool(np.all(run.ierr[0] == 0) and np.all(run.relres[0] <= 1e-8))
    return dict(workload=f"plate, {len(nef)} failed elements (notch tip + free corner), "
                         f"{len(ev.ndfail)} orphan node(s), next slab with {R} cases",
                nefail=int(len(nef)), ndfail=int(len(ev.ndfail)),
                ms={k: 1e3 * v for k, v in t.items()},
                host_edit_ms={k: 1e3 * v for k, v in ev.timings.items()},
                solve_converged=ok, its_max=int(np.max(run.its[0])),
                relres_max=float(np.max(run.relres[0])))


def reassembly_leg(prob, reps=3):
    """Per-slab re-assembly of the plate on the device, as the driver repeats it every
    slab for the temperature-dependent material (main3d.f90:752-783): element matrices
    and global K, M with the clamp (xt_asm_run), new values into the operator
    (xt_op_set_values) and the band refactorization (xt_prec_refactor_band), for a new
    E and nu at every Gauss point.  The host restatement of the same assembly (numpy,
    vectorised over elements) is timed beside it."""
    import torch
    from xtfem_amd import device, fem
    mesh = prob.mesh
    t0 = time.perf_counter()
    asm = device.Assembly(mesh, prob.bc)
    t_sym = time.perf_counter() - t0
    K, M = asm.run(2.0e5, 0.3, 7.85e-9)
    Kp = (asm.K_pattern[0], asm.K_pattern[1], K.cpu().numpy())
    Mp = (asm.M_pattern[0], asm.M_pattern[1], M.cpu().numpy())
    op = device.KronOperator([Kp, Mp], q=6)
    pr = device.BandPreconditioner(Kp)
    rng = np.random.default_rng(SEED + 7)
    t = {"assembly": [], "operator_update": [], "band_refactor": []}
    for _ in range(reps):
        ym = torch.as_tensor(rng.uniform(1.9e5, 2.0e5, 8 * mesh.ne), device="cuda")
        nu = torch.as_tensor(rng.uniform(0.29, 0.31, 8 * mesh.ne), device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K, M = asm.run(ym, nu, 7.85e-9)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        op.set_values(K, M)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        pr.refactor(asm.K_rowptr, asm.K_col, K)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        t["assembly"].append(t1 - t0)
        t["operator_update"].append(t2 - t1)
        t["band_refactor"].append(t3 - t2)
    t0 = time.perf_counter()
    fem.assemble(mesh, ym.cpu().numpy(), nu.cpu().numpy(), 7.85e-9)
    t_host = time.perf_counter() - t0
    ms = {k: 1e3 * float(np.median(v)) for k, v in t.items()}
    return dict(workload=f"plate, {mesh.ne} C3D8 elements, K nnz {len(asm.K_pattern[1])}, "
                         f"new E, nu per Gauss point", ms=ms,
                total_ms=sum(ms.values()), symbolic_setup_s=t_sym,
                host_assembly_ms=1e3 * t_host)


def thermo_leg(prob, nslab=4):
    """The thermo-mechanical slab loop on the plate (main3d.f90:699-840): per slab a
    thermal slab solve (q = 3, exact banded factor of K_t), material update and thermal
    load, re-assembly of K and M, operator update and band refactor, and the mechanical
    slab solve with the thermal load on the right-hand side.  Temperatures fixed on
    y = 0 and y = ny (the reference's z = 0 and z = le), heat flux 2 W/mm^2 lumped on the
    side faces x = 0 and x = nx (its y = +-35), traction on y = ny, slabs of 0.5 s at
    20 Hz.  One warm-up slab, then ``nslab`` timed slabs per solver mode; the host
    restatement of the per-slab assembly work (material update, thermal load, K and M)
    is timed beside it."""
    import torch
    from xtfem_amd import fem, thermo
    mesh = prob.mesh
    nx, ny, nz = (int(round(mesh.xcoord[:, k].max())) for k in range(3))
    tfixed = fem.fixed_temperature_mask(mesh, 1, [0.0, float(ny)])
    rext_t = fem.heat_flux_load(mesh, 0, 0.0, [0.0, float(nx)], 2.0 * ny * nz)
    rext_m = prob.amp[0] * fem.traction_load(mesh, 1, float(ny), 1, float(nx * nz))
    out = {}
    for mode, name, shift in ((0, "fp64_reference_prec", 0.0), (0, "fp64", "auto"),
                              (1, "fp32_ir", "auto")):
        run = thermo.ThermoMechanicalSlabs(mesh, prob.bc, tfixed, rext_m, rext_t, 0.5,
                                           2 * np.pi * 20.0, mode=mode, thermal_shift=shift,
                                           profile=True)
        run.step()
        torch.cuda.synchronize()
        wall, its, phases = [], [], []
        for _ in range(nslab):
            t0 = time.perf_counter()
            its.append(run.step())
            torch.cuda.synchronize()
            wall.append(time.perf_counter() - t0)
            phases.append(run.phase_ms())
        ph = {k: float(np.median([p[k] for p in phases])) for k in run.PHASES}
        T = run.temperature()
        out[name] = dict(ms_per_slab=1e3 * float(np.median(wall)), phases_ms=ph,
                         thermal_shift=run.thermal_shift,
                         its_thermal=[i[0] for i in its], its_mech=[i[1] for i in its],
                         T_max=float(T.max()), T_min=float(T.min()),
                         relres_t=run.last["relres_t"], relres_m=run.last["relres_m"])
        del run
    T = np.zeros(mesh.nn)   # host restatement of the per-slab assembly work
    T[:] = 20.0 + 10.0 * np.sin(mesh.xcoord[:, 0] / nx * np.pi)
    t0 = time.perf_counter()
    ym, nu = fem.material_update(mesh, T, 0.0, thermo.YM0, thermo.NU0)
    fem.thermal_load(mesh, T, 0.0, fem.ALPHA_REF, ym, nu)
    fem.assemble(mesh, ym, nu, thermo.RHO)
    out["host_assembly_ms"] = 1e3 * (time.perf_counter() - t0)
    out["workload"] = (f"plate, {mesh.ne} C3D8 elements, thermal {mesh.nn} DOFs x 3, "
                       f"mechanical {prob.p} DOFs x 6, one load case")
    return out


def driver_panel_leg(nslab=3):
    """The coupled loop on the reference's own panel mesh (panel_7634_3749: 7634 nodes,
    3749 C3D8, -50..50 x -35..35 x 0..130 mm, renumbered by reverse Cuthill-McKee) with
    main3d.f90's boundary conditions, loads and material, slabs of 0.5 s at 20 Hz."""
    import torch
    from xtfem_amd import fem, thermo
    with np.load(os.path.join(ROOT, "tests", "golden", "panel_7634_3749.npz"),
                 allow_pickle=False) as z:
        mesh = fem.Mesh(z["xcoord"], z["xconn"])
    bw0 = fem.node_bandwidth(mesh)
    mesh = fem.renumber(mesh, fem.rcm_order(mesh))
    run = thermo.ThermoMechanicalSlabs(mesh, *thermo.driver_conditions(mesh), 0.5,
                                       2 * np.pi * 20.0, profile=True)
    run.step()
    torch.cuda.synchronize()
    wall, its, phases = [], [], []
    for _ in range(nslab):
        t0 = time.perf_counter()
        its.append(run.step())
        torch.cuda.synchronize()
        wall.append(time.perf_counter() - t0)
        phases.append(run.phase_ms())
    T = run.temperature()
    return dict(workload=f"reference panel mesh, {mesh.nn} nodes, {mesh.ne} C3D8, node "
                         f"half-bandwidth {bw0} -> {fem.node_bandwidth(mesh)} after RCM, "
                         f"band blocks {run.pr_m.block} (mech) / {run.pr_t.block} (thermal)",
                ms_per_slab=1e3 * float(np.median(wall)),
                phases_ms={k: float(np.median([p[k] for p in phases])) for k in run.PHASES},
                its_thermal=[i[0] for i in its], its_mech=[i[1] for i in its],
                T_max=float(T.max()), relres_m=run.last["relres_m"])


_CPU = {}


def _cpu_worker(widx):
    """One host core of the CPU baseline: the oracle's FP64 kronpgmres
    (kronecker.f90:181-365) on consecutive slabs of load case widx mod R (jump term, RHS,
    solve to 1e-8, preconditioner A_r^-1 (x) K^-1 with the exact factor), until the time
    budget is spent (at least one slab).  BLAS runs single-threaded in the worker."""
    from threadpoolctl import threadpool_limits
    from oracle import oracle
    from xtfem_amd import slab
    prob, lf, budget = _CPU["prob"], _CPU["factor"], _CPU["budget"]
    r = widx % len(prob.amp)
    p = prob.p
    fixed = np.tile(prob.bc, 6)
    x = np.zeros(6 * p)
    n, its_all = 0, []
    with threadpool_limits(1):
        t0 = time.perf_counter()
        while n == 0 or time.perf_counter() - t0 < budget:
            jump = oracle.kronspmv(prob.J1[r], prob.J2[r], prob.Kb, prob.Mb, x)
            rhs = slab.slab_rhs(p, prob.coef(n, np.array([r])), None, prob.rext[r], jump=jump,
                                fixed=fixed)
            x, ie, its, _ = oracle.kronpgmres(prob.A[r], prob.B[r], prob.Kb, prob.Mb, lf, rhs,
                                              im=30, eps=1e-8, maxits=300)
            its_all.append(int(its))
            n += 1
        t = time.perf_counter() - t0
    return n, t, its_all


def cpu_baseline(prob, budget_s=45.0):
    """The oracle on the host, one process per host core (the reference runs one
    single-threaded Kokkos instance per MPI rank, kokkos_host.cpp:17): each worker solves
    consecutive slabs of its own load case with the oracle's FP64 kronpgmres and the exact
    factor of K (LAPACK band Cholesky, computed once before the workers fork and shared;
    its time, on all cores, is reported and excluded, as the GPU factorization is).
    value = slabs finished by all workers / wall time of the parallel section."""
    import multiprocessing as mp
    from oracle import oracle
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nw = max(1, min(ncores, 16))
    t0 = time.perf_counter()
    lf = oracle.LapackBandFactor(prob.Kb)
    t_fact = time.perf_counter() - t0
    _CPU.update(prob=prob, factor=lf, budget=budget_s)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(nw) as pool:
        res = pool.map(_cpu_worker, range(nw))
    wall = time.perf_counter() - t0
    nsl = sum(r[0] for r in res)
    its = [i for r in res for i in r[2]]
    return dict(value=nsl / wall, unit="slab solves/s", cores=nw, kind="port",
                sample=f"{nw} worker processes (one per host core, single-threaded BLAS), each "
                       f"on consecutive slabs of its own load case for about {budget_s:.0f} s: "
                       f"jump term, RHS and the oracle's FP64 GMRES(30) to 1e-8 with "
                       f"A_r^-1 (x) K^-1, K^-1 the exact band Cholesky factor (half-bandwidth "
                       f"{lf.w}; LAPACK factorization {t_fact:.1f} s on {ncores} cores, "
                       f"excluded); {nsl} slabs in {wall:.1f} s, GMRES its min {min(its)} max "
                       f"{max(its)}", nproc=ncores, factor_s=t_fact, slabs=nsl, wall_s=wall)


def cpu_only(args):
    """--cpu-only: the CPU baseline in a process that never touches the GPU (its workers
    fork from it).  The parent passes the workload's matrices in a directory."""
    R = args.cases
    if args.workload == "panel":
        prob = PanelProblem.__new__(PanelProblem)
        grid = tuple(args.grid) if args.grid else (240, 200, 2)
        prob.grid = grid
        prob.mesh, prob.bc = panel_mesh(*grid)
        prob.p = 3 * prob.mesh.nn
        from xtfem_amd import fem, slab
        prob.amp, prob.freq, prob.dt = load_cases(R, SEED + 3)
        prob.w = 2 * np.pi * prob.freq
        c = slab.stcoeff_m(0.0, prob.dt, prob.w)
        prob.A, prob.B = c["k"] + c["ktn"], c["m"] + c["mtn"]
        prob.J1, prob.J2 = c["ktn1"], c["mtn1"]
        unit = fem.traction_load(prob.mesh, 1, float(grid[1]), 1, float(grid[0] * grid[2]))
        prob.rext = np.tile(unit, (R, 1))
    else:
        grid = args.grid or [100, 46, 4]
        prob = Problem(*grid, args.notch, R, args.cases_kind)
    if args.cpu_matrices:
        d = args.cpu_matrices
        ld = lambda n: np.load(os.path.join(d, n + ".npy"), allow_pickle=False)
        prob.Kb = (ld("kia"), ld("kja"), ld("kv"))
        prob.Mb = (ld("mia"), ld("mja"), ld("mv"))
    elif not hasattr(prob, "Kb"):
        from xtfem_amd import fem
        K, M, *_ = fem.assemble(prob.mesh, 2.0e5, 0.3, 7.85e-9)
        prob.Kb, prob.Mb = fem.applybc(K, prob.bc), fem.applybc(M, prob.bc)
    out = cpu_baseline(prob, args.cpu_budget)
    out["damage"] = cpu_damage_sample()
    print(json.dumps(out))


def run_cpu_baseline(args, prob):
    """The CPU baseline in a child process (no GPU in it, so its workers may fork), with
    the workload's matrices handed over through a temporary directory."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        for name, arr in zip(("kia", "kja", "kv"), prob.Kb):
            np.save(os.path.join(d, name + ".npy"), np.asarray(arr))
        for name, arr in zip(("mia", "mja", "mv"), prob.Mb):
            np.save(os.path.join(d, name + ".npy"), np.asarray(arr))
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-only", "--workload",
               args.workload, "--cases", str(args.cases), "--cpu-budget", str(args.cpu_budget),
               "--cpu-matrices", d, "--notch", str(args.notch), "--cases-kind", args.cases_kind]
        if args.grid:
            cmd += ["--grid", *map(str, args.grid)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            print(r.stderr[-2000:], file=sys.stderr)
            return None
        return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_rod_sample(slabs=5):
    """The oracle's FP64 kronpgmres with its ILUT(20, 1e-4) factors on the rod, one core."""
    from oracle import oracle
    from xtfem_amd import slab
    prob = RodProblem()
    p = prob.p
    prec = oracle.ilut(prob.Kb, 20, 1e-4)[:3]
    fixed = np.tile(prob.bc, 6)
    x = np.zeros(6 * p)
    t0 = time.perf_counter()
    its_all = []
    for n in range(slabs):
        jump = oracle.kronspmv(prob.J1[0], prob.J2[0], prob.Kb, prob.Mb, x)
        rhs = slab.slab_rhs(p, prob.coef(n, np.array([0])), None, prob.rext[0], jump=jump,
                            fixed=fixed)
        x, ie, its, _ = oracle.kronpgmres(prob.A[0], prob.B[0], prob.Kb, prob.Mb, prec, rhs,
                                          im=30, eps=1e-8, maxits=500)
        its_all.append(its)
    t = time.perf_counter() - t0
    return dict(slabs_per_s=slabs / t, cores=1, kind="port",
                sample=f"{slabs} slabs, oracle FP64 kronpgmres + ILUT(20, 1e-4), its {its_all}")


def cpu_damage_sample(ne_s=1000, nip=2560):
    from oracle import oracle
    from xtfem_amd import slab
    dp = damage_problem()
    mesh, Nx, Ny, Nz, sol = dp["mesh"], dp["Nx"], dp["Ny"], dp["Nz"], dp["sol"]
    ngp = 8 * ne_s
    nodes = np.repeat(mesh.xconn[:ne_s], 8, axis=0).ravel()
    Nt = slab.damage_nt(np.array([0.0, 0.25, 0.5]), 2 * np.pi * 20.0, nip).ravel()
    st = {k: np.zeros(6 * ngp) for k in ("eps_old", "eps_pl", "sig_eff", "sig_back")}
    st.update({k: np.zeros(ngp) for k in ("p", "D", "ws", "del_po", "sig_d")})
    t0 = time.perf_counter()
    oracle.kokkos_host(ne_s, mesh.nn, nip, nodes, np.ravel(Nx)[:8 * ngp], np.ravel(Ny)[:8 * ngp],
                       np.ravel(Nz)[:8 * ngp], Nt, sol, np.full(ngp, 1.97e5),
                       np.full(ngp, 0.3), st)
    t = time.perf_counter() - t0
    return dict(gp_steps_per_s=ngp * nip / t, cores=1, kind="port",
                sample=f"oracle kokkos_host, {ngp} Gauss points x {nip} temporal points, "
                       f"{t:.1f} s on one host core")


# rocprofv3 --pmc summaries (FETCH_SIZE / WRITE_SIZE per launch, calibrated on a plain
# copy), per workload; none is committed for the current kernels (the counter runs
# crashed in the profiler's initialisation on this image), so traffic reads null
PMC_SUMMARY = {"panel": os.path.join(ROOT, "profiles", "band_sweep_panel_32cases_pmc.md"),
               "plate": os.path.join(ROOT, "profiles", "band_sweep_notched_plate_32cases_pmc.md")}


def pmc_traffic(name, workload):
    """HBM bytes per launch of kernel `name` in `workload` from the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration), or
    None when that summary does not give it."""
    import re
    path = PMC_SUMMARY.get(workload)
    if not path or not os.path.exists(path):
        return None
    text = open(path).read()
    m = re.search(r"^traffic_bytes_per_launch\[" + re.escape(name) + r"\]:\s*([0-9.eE+]+)",
                  text, re.M)
    return float(m.group(1)) if m else None


def roofline(name, k, prob, pr, workload):
    """Roofline entry of the dominant kernel, on its algorithmic work (xt_prof_read_alg):
    matrix-core bound when the algorithmic intensity is above the ridge of its precision
    (FP32 157.3, FP64 78.6 TFLOP/s over 8 TB/s), else HBM bound.  For the factor sweeps
    the algorithmic flops are those of the unpaired block recurrences (the pairing's extra
    products are reported as performed); the banded minimum 4 p w c flops per apply (two
    triangular sweeps over half-bandwidth w) is given beside it."""
    prec = "f64" if name.endswith("_f64") else "f32"
    peak = PEAK_MFMA_TFLOPS[prec]
    intensity = k["flops"] / k["bytes"]
    out = {"kernel": name, "traffic": pmc_traffic(name, workload),
           "algorithmic_bytes_per_launch": k["bytes"] / k["launches"],
           "algorithmic_flops_per_launch": k["flops"] / k["launches"],
           "performed_flops_per_launch": k["performed_flops"] / k["launches"],
           "avg_launch_us": k["avg_us"], "intensity_flop_per_byte": intensity}
    if intensity > peak * 1e12 / (PEAK_HBM_GBS * 1e9):
        tf = k["flops"] / (k["ms"] * 1e-3) / 1e12
        out.update(bound="mfma", achieved=tf, peak=peak, unit="TFLOP/s", frac=tf / peak,
                   performed_tflops=k["performed_flops"] / (k["ms"] * 1e-3) / 1e12)
        if name.startswith("band_sweep"):
            out["banded_minimum_flops_per_apply_c192"] = 4.0 * prob.p * pr.bandwidth * 192
    else:
        out.update(bound="hbm", achieved=k["gbs"], peak=PEAK_HBM_GBS, unit="GB/s",
                   frac=k["gbs"] / PEAK_HBM_GBS)
    return out


def copy_probe():
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 1 << 28                                     # 2^28 floats = 1 GiB each way
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))


def setup_solver(prob, args, fp32_only):
    """Operator and preconditioner of the workload: the exact banded factor of K and the
    shifted factors (K + s M) covering the cases' dominant temporal eigenvalues above
    0.05 of the structure's lowest K^-1 M eigenvalue (device power iteration)."""
    import torch
    from xtfem_amd import device
    from xtfem_amd import tune
    t0 = time.perf_counter()
    # the factors' pairing levels (built-in unless --band-levels; the sweeps use them per
    # apply by column count: at 192 columns the panel's FP64 chain steps fill the GPU
    # unpaired, 27.6 vs 31.0 ms, at a few columns pairing wins, 11.0 vs 11.9 ms)
    tune.set_knob("band_levels", args.band_levels)
    op = device.KronOperator([prob.Kb, prob.Mb], q=6)
    pr0 = device.BandPreconditioner(prob.Kb, block=args.block)
    mu = device.lowest_mode(pr0, prob.Mb, prob.bc)
    del pr0
    torch.cuda.synchronize()
    lam = device.dominant_temporal_mode(prob.A, prob.B)
    shifts = [] if args.shift_ratio <= 0 else device.shifts_for_cases(lam, mu, args.shift_ratio)
    t1 = time.perf_counter()
    pr = device.BandPreconditioner(prob.Kb, block=args.block, M=prob.Mb, shifts=shifts,
                                   shifted_fp32_only=fp32_only)
    torch.cuda.synchronize()
    tune.set_knob("band_levels", 0)
    info = dict(mu_min=mu, lambda_min=float(np.nanmin(lam)), lambda_max=float(np.nanmax(lam)),
                shifts=[float(v) for v in shifts], shifted_fp32_only=fp32_only,
                bandwidth=pr.bandwidth, blocks=pr.nblocks, block=pr.block,
                setup_s=dict(operator_and_mu=t1 - t0, factors=time.perf_counter() - t1))
    return op, pr, info


def accuracy_check(prob, op, pr, xtol):
    """Slab 0 of every case (x_prev = 0): FP32-IR, FP64 GMRES at 1e-8 and an FP64 solve to
    1e-12; relative 2-norm displacement error of the first two against the third."""
    import torch
    from xtfem_amd import device
    R = len(prob.amp)
    runs = {}
    for name, mode, eps, xt in (("ref", 0, 1e-12, 0.0), ("fp32_ir", 1, 1e-8, xtol),
                                ("fp64", 0, 1e-8, 0.0)):
        run = Runner(prob, op, pr, np.arange(R), mode=mode, nsteps=1, xtol=xt)
        run.eps = eps
        run.step()
        runs[name] = run
    x = {k: device.from_interleaved(v.xprev, prob.p, 6, R) for k, v in runs.items()}
    err = {k: np.linalg.norm(x[k] - x["ref"], axis=1) / np.linalg.norm(x["ref"], axis=1)
           for k in ("fp32_ir", "fp64")}
    return dict(rel_err_fp32_ir_max=float(err["fp32_ir"].max()),
                rel_err_fp64_max=float(err["fp64"].max()),
                ref_relres_max=float(np.max(runs["ref"].relres[0])),
                ref_its_max=int(np.max(runs["ref"].its[0])),
                fp32_ir_its=runs["fp32_ir"].its[0].tolist(),
                fp64_its=runs["fp64"].its[0].tolist())


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(independent load-case groups or meshes would shard across GPUs).",
              file=sys.stderr)
        sys.exit(2)
    if args.cpu_only:
        cpu_only(args)
        return
    import torch
    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.rod_only:
        print(json.dumps(rod_leg(args.steps)))
        return
    if args.damage_only:
        print(json.dumps(damage_leg(args.damage_nip, slabs=args.steps)))
        return
    if args.fatigue_only:
        print(json.dumps(fatigue_leg(args.fatigue_blocks, args.damage_nip)))
        return
    R, K, W = args.cases, args.steps, args.warmup
    if args.workload == "multimesh":
        print(json.dumps(multimesh_leg(K=K)))
        return
    if args.thermo_only:
        grid = args.grid or [100, 46, 4]
        print(json.dumps(thermo_leg(Problem(*grid, args.notch, 1), nslab=args.steps)))
        return
    t0 = time.perf_counter()
    if args.workload == "panel":
        prob = PanelProblem(R, tuple(args.grid) if args.grid else (240, 200, 2))
        fp32_only = True      # FP64 shifted factors of the panel would not fit beside the rest
    else:
        prob = Problem(*(args.grid or [100, 46, 4]), args.notch, R, args.cases_kind)
    if args.spmm_only:
        from xtfem_amd import device
        op = device.KronOperator([prob.Kb, prob.Mb], q=6)
        print(json.dumps(dict(workload=prob.describe(), kron_spmm=spmm_legs(op, prob.p))))
        return
        fp32_only = False
    t_asm = time.perf_counter() - t0
    op, pr, sinfo = setup_solver(prob, args, fp32_only)
    print(f"# setup: p={prob.p} assembly {t_asm:.1f}s, {sinfo}", file=sys.stderr, flush=True)

    nsteps = W + K
    res = {}
    run32 = Runner(prob, op, pr, np.arange(R), mode=1, nsteps=2 * nsteps, xtol=args.xtol)
    for _ in range(W):
        run32.step()
    if args.profile_only:
        # the batched GMRES-IR steps with per-launch HIP events only (for a kernel trace of
        # this command: its per-kernel averages against the events' averages printed here)
        prof = kernel_profile(run32, K, mask=0b10011)
        dom = max(prof, key=lambda k: prof[k]["ms"])
        print(json.dumps(dict(workload=prob.describe(), steps=K, warmup=W,
                              its=run32.its[-1].tolist(),
                              iterations_by_sweep_precision=np.array(run32.it_prec[W:]).sum(
                                  axis=0).tolist(),
                              kernels=prof, roofline=roofline(dom, prof[dom], prob, pr,
                                                              args.workload))))
        return
    t32 = timed(run32, K)
    it_prec = np.array(run32.it_prec[W:]).sum(axis=0)      # (FP32, FP64) sweep iterations
    its32 = np.array(run32.its[W:])
    rr32 = np.array(run32.relres[W:])
    ok32 = bool(np.all(np.array(run32.ierr[W:]) == 0) and np.all(rr32 <= 1e-8))
    print(f"# fp32-ir batched: {K} steps x {R} cases in {t32:.3f}s, its max {its32.max()}, "
          f"relres max {rr32.max():.2e}; its per case (last slab) {run32.its[-1].tolist()}",
          file=sys.stderr, flush=True)
    if not args.no_fp64:
        # plain FP64 GPU GMRES, the reference's preconditioner A^-1 (x) K^-1 (the shifted
        # factors of the panel are FP32 only), the whole batch in one solve
        runb = Runner(prob, op, pr, np.arange(R), mode=0, nsteps=nsteps)
        for _ in range(W):
            runb.step()
        tb = timed(runb, K)
        res["fp64_batched_slabs_per_s"] = K * R / tb
        res["fp64_batched_its_max"] = int(np.max(runb.its[W:]))
        res["fp64_batched_relres_max"] = float(np.max(runb.relres[W:]))
        res["fp64_batched_all_converged"] = bool(np.all(np.array(runb.ierr[W:]) == 0))
        print(f"# fp64 batched: {K} steps x {R} cases in {tb:.3f}s, its max "
              f"{np.max(runb.its[W:])}", file=sys.stderr, flush=True)
        del runb
        # the same to 1e-12: the FP64 run whose displacements meet 1e-6 where the slab
        # operator is ill-conditioned (the accuracy check below gives both errors)
        runt = Runner(prob, op, pr, np.arange(R), mode=0, nsteps=nsteps)
        runt.eps = 1e-12
        for _ in range(W):
            runt.step()
        tt = timed(runt, K)
        res["fp64_batched_eps1e-12_slabs_per_s"] = K * R / tt
        res["fp64_batched_eps1e-12_its_max"] = int(np.max(runt.its[W:]))
        print(f"# fp64 batched to 1e-12: {K} steps x {R} cases in {tt:.3f}s, its max "
              f"{np.max(runt.its[W:])}", file=sys.stderr, flush=True)
        del runt
        # the same, one load case per solve
        runs = [Runner(prob, op, pr, np.array([r]), mode=0, nsteps=nsteps) for r in range(R)]
        for _ in range(W):
            for rn in runs:
                rn.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            for rn in runs:
                rn.step()
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        its64 = np.array([rn.its[W:] for rn in runs])
        res["fp64_single_slabs_per_s"] = K * R / t64
        res["fp64_single_its_mean"] = float(its64.mean())
        res["fp64_single_relres_max"] = float(np.max([rn.relres[W:] for rn in runs]))
        print(f"# fp64 one case per solve: {K}x{R} slabs in {t64:.3f}s, its mean "
              f"{its64.mean():.1f}", file=sys.stderr, flush=True)
        del runs
        # the same to 1e-12: one case per solve at the headline's accuracy
        runs = [Runner(prob, op, pr, np.array([r]), mode=0, nsteps=nsteps) for r in range(R)]
        for rn in runs:
            rn.eps = 1e-12
        for _ in range(W):
            for rn in runs:
                rn.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            for rn in runs:
                rn.step()
        torch.cuda.synchronize()
        t64t = time.perf_counter() - t0
        res["fp64_single_eps1e-12_slabs_per_s"] = K * R / t64t
        res["fp64_single_eps1e-12_its_mean"] = float(np.mean([rn.its[W:] for rn in runs]))
        print(f"# fp64 one case per solve to 1e-12: {K}x{R} slabs in {t64t:.3f}s",
              file=sys.stderr, flush=True)
        del runs
        res["precision_ratio"] = (K * R / t32) / res["fp64_batched_slabs_per_s"]
        res["precision_ratio_equal_accuracy"] = ((K * R / t32) /
                                                 res["fp64_batched_eps1e-12_slabs_per_s"])
        res["batching_ratio"] = res["fp64_batched_slabs_per_s"] / res["fp64_single_slabs_per_s"]
        res["speedup_vs_fp64_single"] = (K * R / t32) / res["fp64_single_slabs_per_s"]
        res["speedup_vs_fp64_single_equal_accuracy"] = ((K * R / t32) /
                                                        res["fp64_single_eps1e-12_slabs_per_s"])
    res["accuracy_slab0"] = accuracy_check(prob, op, pr, args.xtol)
    print(f"# accuracy (slab 0 vs FP64 at 1e-12): {res['accuracy_slab0']}", file=sys.stderr,
          flush=True)
    spmm = spmm_legs(op, prob.p)
    print(f"# kron-spmm fp32: R=1 {spmm['R1']['gbs']:.0f} GB/s ({spmm['R1']['frac']:.2f}), "
          f"R=32 {spmm['R32']['gbs']:.0f} GB/s ({spmm['R32']['frac']:.2f}; "
          f"{spmm['R32']['bound']}-bound, {spmm['R32']['attainable_frac']:.2f} of that roof)",
          file=sys.stderr,
          flush=True)
    # per-kernel durations: the batched FP32-IR steps once more, every launch of the
    # operator and sweep kernels bracketed by HIP events on the launching stream
    prof = kernel_profile(run32, K, mask=0b10011)
    dom_name = max(prof, key=lambda k: prof[k]["ms"])
    roof = roofline(dom_name, prof[dom_name], prob, pr, args.workload)
    value = K * R / t32
    ms_step = 1e3 * t32 / K
    share32 = float(it_prec[0] / max(1, it_prec.sum()))
    # the arithmetic of the timed steps: GMRES-IR runs FP32 cycles where they pay and FP64
    # cycles elsewhere (the solver probes FP32 and remembers when it did not pay)
    dtype = ("f32 sweeps/operator in f64 GMRES-IR" if share32 >= 0.9 else
             "f64 (GMRES-IR: FP32 cycles probed, not profitable for this K)" if share32 <= 0.1
             else "f32+f64 (GMRES-IR, mixed cycles)")
    res["iterations_by_sweep_precision"] = {"f32": int(it_prec[0]), "f64": int(it_prec[1]),
                                            "f32_share": share32}
    nblocks, block = pr.nblocks, pr.block
    del run32, op, pr                  # the extra legs need the memory
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    legs = set(args.legs.split(",")) - {""}
    if "all" in legs:
        legs = {"rod", "reassembly", "thermo", "driver_panel", "deletion", "damage", "fatigue",
                "multimesh"}
    plate = prob if isinstance(prob, Problem) else None
    if legs & {"reassembly", "thermo", "deletion"} and plate is None:
        plate = Problem(100, 46, 4, 20, min(R, 8), "plate")
    if "rod" in legs:
        res["rod_ilut"] = rod_leg(K)
    if "reassembly" in legs:
        res["reassembly"] = reassembly_leg(plate)
    if "thermo" in legs:
        res["thermo"] = thermo_leg(plate)
    if "driver_panel" in legs:
        res["driver_panel"] = driver_panel_leg()
    if "deletion" in legs:
        res["deletion"] = deletion_leg(plate)
    if "damage" in legs:
        res["damage"] = damage_leg(args.damage_nip)
    if "multimesh" in legs:
        res["multimesh"] = multimesh_leg()
        print(f"# multimesh: {res['multimesh']}", file=sys.stderr, flush=True)
    if "fatigue" in legs:
        res["fatigue"] = fatigue_leg(args.fatigue_blocks, args.damage_nip)
        print(f"# fatigue: {res['fatigue']}", file=sys.stderr, flush=True)
    out = {
        "metric": "slab solves/s at rel.res 1e-8 (FP32-IR batched vs FP64 GPU); Kron-SpMM GB/s",
        "value": value,
        "unit": "slab solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (structured mesh, seeded load cases)",
        "config": {"workload": prob.describe() + f"; GMRES(30) eps 1e-8, forward-error test "
                               f"{args.xtol:g}; exact banded factors ({nblocks} blocks of "
                               f"{block}), {len(sinfo['shifts'])} shifted",
                   "cases": R, "p": prob.p},
        "all_converged": ok32,
        "relres_max": float(rr32.max()),
        "its_max": int(its32.max()),
        "its_mean": float(its32.mean()),
        **res,
        "solver_setup": sinfo,
        "kron_spmm": spmm,
        "kernels": prof,
        "roofline": roof,
        "setup_s": {"assembly": t_asm},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline(args, prob)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
