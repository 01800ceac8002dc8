"""Write the golden fixtures of tests/golden from the reference's own code.

Run in a container where /root/reference is mounted and ``make -C oracle ref`` has
built oracle/_ref/libxtfem_ref.so (the reference's Fortran, compiled in place):

    python tests/golden/make_golden.py

Outputs (numpy .npz, no pickles):
  rod_325_192.npz   the mesh inp/prismaticrod-325-192.inp (input data: coordinates,
                    connectivity)
  golden_ref.npz    input/output vectors of the compiled reference routines
                    (stcoeff_m, stcoeff_t, getfextc_m, mpiglbmtx_m+tricsr,
                    csrapplybc, ilut, lusol, kronspmv, kronspgemv, kronlusol,
                    kronpgmres, getdfinal, chkelesta, the element-deletion block) on
                    seeded inputs.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from xtfem_amd import fem  # noqa: E402

SEED = 20261202
INP = "/root/reference/inp/prismaticrod-325-192.inp"
RHO_REF = float(np.float32(7.860e-6))   # main3d.f90:154, default-real literal
YM_REF, NU_REF = 1.97e5, 0.3             # main3d.f90:152-153


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def strain_history(nsteps, amp, dirs, ph):
    """Smooth cyclic strain path (6 Voigt components, 3 cycles) of amplitude amp.
    The damage tests rebuild the same paths from (amp, dirs, ph)."""
    t = np.arange(nsteps) / nsteps
    return amp * dirs[None, :] * np.sin(2 * np.pi * 3 * t[:, None] + ph[None, :])


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref/libxtfem_ref.so missing: run `make -C oracle ref`")
    mesh = fem.read_inp(INP)
    np.savez_compressed(os.path.join(HERE, "rod_325_192.npz"), xcoord=mesh.xcoord,
                        xconn=mesh.xconn)
    # the driver's own panel mesh (the geometry main3d.f90's boundary conditions are
    # written for: -50..50 x -35..35 x 0..130 mm), kept as the reader returns it
    pan = fem.read_inp(os.path.join(os.path.dirname(INP), "panel_7634_3749.inp"))
    np.savez_compressed(os.path.join(HERE, "panel_7634_3749.npz"), xcoord=pan.xcoord,
                        xconn=pan.xconn)
    rng = np.random.default_rng(SEED)
    g = {}
    # temporal matrices and load coefficients (stcoeff.f90, getfextc.f90)
    pts = np.array([[0.0, 0.5, 2 * np.pi * 20], [0.0, 0.05, 2 * np.pi * 20],
                    [0.3, 0.5, 2 * np.pi * 50], [0.0, 0.1, 2 * np.pi * 97.3]])
    g["st_pts"] = pts
    for i, (t1, dt, w) in enumerate(pts):
        c = ref.stcoeff_m(t1, dt, w)
        for k, v in c.items():
            g[f"st{i}_{k}"] = v
        f, h = ref.getfextc_m(t1, w, dt)
        g[f"fx{i}_fextc"] = f
        g[f"fx{i}_fint"] = h
    ct = ref.stcoeff_t(0.0, 0.5)
    for k, v in ct.items():
        g[f"stt_{k}"] = v
    # assembly of the rod (mpiglbmtx_m + tricsr) and 1-0 BCs clamping x = 0
    K, M, Nx, Ny, Nz = ref.assemble_m(mesh.xcoord, mesh.xconn, YM_REF, NU_REF, RHO_REF)
    g["asm_K_digest"] = np.array(digest(*K))
    g["asm_M_digest"] = np.array(digest(*M))
    g["asm_Nxyz_digest"] = np.array(digest(Nx.T, Ny.T, Nz.T))
    g["asm_K_nnz"] = np.array(len(K[2]))
    bc = fem.clamp_mask(mesh, 0, [0.0])
    Kb = ref.csrapplybc(K, bc)
    Mb = ref.csrapplybc(M, bc)
    g["bc_mask"] = bc
    g["Kb_digest"] = np.array(digest(*Kb))
    g["Mb_digest"] = np.array(digest(*Mb))
    # ILUT on the rod, lfil=20, droptol=1e-4 (the rod benchmark settings) and a tighter variant
    alu, jlu, ju, ierr = ref.ilut(Kb, 20, 1e-4)
    assert ierr == 0
    g["ilut20_digest"] = np.array(digest(alu, jlu, ju))
    g["ilut20_nnz"] = np.array(len(alu))
    alu5, jlu5, ju5, ierr5 = ref.ilut(Kb, 5, 1e-3)
    assert ierr5 == 0
    g["ilut5_digest"] = np.array(digest(alu5, jlu5, ju5))
    # triangular solve, Kronecker ops, preconditioner, GMRES
    p = len(Kb[0]) - 1
    y = rng.standard_normal(p)
    g["lusol_y"] = y
    g["lusol_x"] = ref.lusol(alu, jlu, ju, y)
    c0 = ref.stcoeff_m(0.0, 0.5, 2 * np.pi * 20)
    A = c0["k"] + c0["ktn"]
    B = c0["m"] + c0["mtn"]
    g["A"], g["B"] = A, B
    x = rng.standard_normal(6 * p)
    g["kron_x"] = x
    g["kron_y"] = ref.kronspmv(A, B, Kb, Mb, x)
    y0 = rng.standard_normal(6 * p)
    g["gemv_y0"] = y0
    g["gemv_y"] = ref.kronspgemv(A, B, 0.7, -1.3, Kb, Mb, x, y0)
    g["klu_y"] = ref.kronlusol(A, alu, jlu, ju, x)
    # a smooth slab right-hand side: axial traction on x = 3 at 20 Hz (t1 = 0, dt = 0.5)
    rext = fem.traction_load(mesh, 0, 3.0, 0, 70.0 * 1.0)
    f, h = ref.getfextc_m(0.0, 2 * np.pi * 20, 0.5)
    rhs = np.concatenate([f[j] * rext for j in range(6)])
    rhs[np.tile(bc, 6) != 0] = 0.0
    g["gmres_rhs"] = rhs
    for eps, tag in ((1e-8, "e8"), (1e-12, "e12")):
        sol, ierr, log = ref.kronpgmres(A, B, Kb, Mb, (alu, jlu, ju), rhs, im=30, eps=eps,
                                        maxits=3000, log=True)
        assert ierr == 0, ierr
        g[f"gmres_{tag}_sol"] = sol
        g[f"gmres_{tag}_log"] = np.array(log)
    # two-scale damage return map (getdfinal.f) on seeded strain paths
    ngp, nsteps = 24, 1200
    E, nu = YM_REF, NU_REF
    G = E / (2 * (1 + nu))
    cf = E / ((1 + nu) * (1 - 2 * nu))
    dmat = np.zeros((6, 6), order="F")
    cmat = np.zeros((6, 6), order="F")
    for i in range(3):
        dmat[i, i] = cf * (1 - nu)
        dmat[i + 3, i + 3] = cf * (1 - 2 * nu) / 2
        cmat[i, i] = 1 / E
        cmat[i + 3, i + 3] = 2 * (1 + nu) / E
        for j in range(3):
            if i != j:
                dmat[i, j] = cf * nu
                cmat[i, j] = -nu / E
    params = np.zeros((ngp, 13))
    in6, in1, strains, out6, out1 = [], [], [], [], []
    for k in range(ngp):
        lo, hi = ((0.0008, 0.004), (0.01, 0.03), (0.08, 0.15))[k % 3]
        amp = rng.uniform(lo, hi)
        dirs = rng.standard_normal(6)
        dirs /= np.linalg.norm(dirs)
        ph = rng.uniform(0, 2 * np.pi, 6)
        params[k] = np.r_[amp, dirs, ph]
        path = strain_history(nsteps, amp, dirs, ph)
        eo = np.zeros(6); ep = np.zeros(6); se = np.zeros(6); sb = np.zeros(6)
        sc = (0.0, 0.0, 0.0, 0.0, 0.0)
        for s_ in range(nsteps):
            st = np.ascontiguousarray(path[s_])
            rec = s_ % 48 == 47
            if rec:
                in6.append(np.array([eo, ep, se, sb]))
                in1.append(np.array(sc))
                strains.append(st.copy())
            sc = ref.getdfinal(eo, st, ep, *sc[:3], se, sb, *sc[3:], cmat, dmat)
            if rec:
                out6.append(np.array([eo, ep, se, sb]))
                out1.append(np.array(sc))
    # single return-map steps of getdfinal sampled along its own trajectories
    g["dmg_params"] = params
    g["dmg_nsteps"] = np.array(nsteps)
    g["dmg_in6"], g["dmg_in1"], g["dmg_strain"] = np.array(in6), np.array(in1), np.array(strains)
    g["dmg_out6"], g["dmg_out1"] = np.array(out6), np.array(out1)
    # element status (chkelesta)
    D = rng.uniform(0.0, 0.35, 8 * mesh.ne)
    nefail, elesta, gpsta = ref.chkelesta(mesh.nn, mesh.ne, 8, 0.3, D)
    g["chk_D"], g["chk_nefail"], g["chk_elesta"], g["chk_gpsta"] = D, nefail, elesta, gpsta
    # element deletion block (main3d.f90:650-686), two successive events on the rod:
    # a corner element plus five random ones, then every element around an interior
    # node; "del" chains the edited matrices, "delra" starts each event from the
    # re-assembled matrices with the current fixed set, as the driver does
    # (main3d.f90:752-783)
    rng2 = np.random.default_rng(SEED + 1)
    deg = fem.node_degree(mesh).astype(np.int32)
    corner = int(np.flatnonzero(deg == 1)[0]) + 1
    e_corner = int(np.flatnonzero((mesh.xconn == corner).any(axis=1))[0]) + 1
    inner = int(np.flatnonzero(deg == 8)[len(np.flatnonzero(deg == 8)) // 2]) + 1
    ev1 = np.unique(np.r_[e_corner, rng2.choice(mesh.ne, 5, replace=False) + 1]).astype(np.int32)
    ring = np.flatnonzero((mesh.xconn == inner).any(axis=1)) + 1
    ev2 = np.setdiff1d(ring, ev1).astype(np.int32)
    for tag, readd in (("del", False), ("delra", True)):
        Kc, Mc = Kb, Mb
        fixed, necnt = bc.astype(np.int32), deg.copy()
        for i, nef in enumerate((ev1, ev2)):
            if readd and i > 0:
                Kc, Mc = ref.csrapplybc(K, fixed), ref.csrapplybc(M, fixed)
            Kc, Mc, fixed, necnt, ndfail = ref.deletion_block(mesh.xconn, mesh.xcoord, Kc, Mc,
                                                              fixed, necnt, nef, YM_REF, NU_REF)
            g[f"{tag}{i}_nefail"], g[f"{tag}{i}_ndfail"] = nef, ndfail
            g[f"{tag}{i}_fixed"], g[f"{tag}{i}_necnt"] = fixed, necnt
            g[f"{tag}{i}_K_digest"] = np.array(digest(*Kc))
            g[f"{tag}{i}_M_digest"] = np.array(digest(*Mc))
            g[f"{tag}{i}_K_nnz"] = np.array(len(Kc[2]))
    # thermal side: mpiglbmtx_t + tricsr (conductivity, capacity), material_update and
    # the thermal load vv of mpiglbmtx_m for a temperature field on the rod, and the
    # heat-flux coefficients getfextc_t
    rho = RHO_REF
    Kt, Mt = ref.assemble_t(mesh.xcoord, mesh.xconn, 0.1, 452.0, rho)
    g["th_Kt_digest"], g["th_Mt_digest"] = np.array(digest(*Kt)), np.array(digest(*Mt))
    g["th_Kt_nnz"] = np.array(len(Kt[2]))
    rng3 = np.random.default_rng(SEED + 2)
    T = np.concatenate([rng3.uniform(-5.0, 60.0, 2 * mesh.nn),
                        20.0 + 40.0 * np.sin(mesh.xcoord[:, 0] / 40.0)
                        - 8.0 * (mesh.xcoord[:, 1] > 6.0)])
    g["th_T"] = T
    for tag, Tref in (("", 0.0), ("r", 4.0)):
        ym, nu = ref.material_update(mesh.xconn, mesh.nn, T, Tref, YM_REF, NU_REF)
        g[f"th{tag}_ym"], g[f"th{tag}_nu"] = ym, nu
        Kc, Mc, _, _, _, vv = ref.assemble_m(mesh.xcoord, mesh.xconn, YM_REF, NU_REF, rho,
                                             ym=ym, nu=nu, temp=T[2 * mesh.nn:],
                                             alpha=float(np.float32(1.2e-5)), Tref=Tref,
                                             with_vv=True)
        g[f"th{tag}_vv"] = vv
        g[f"th{tag}_K_digest"] = np.array(digest(*Kc))
        g[f"th{tag}_M_digest"] = np.array(digest(*Mc))
    tpts = np.array([[0.0, 0.5], [0.3, 0.05], [12.0, 1.0], [7.5, 0.05]])
    g["fxt_pts"] = tpts
    g["fxt"] = np.array([ref.getfextc_t(t1, dt) for t1, dt in tpts])
    # result files of one output step (mpipostprocess, postprocess.f90:106-210) on a
    # distorted 3 x 2 x 2 block with per-point material, a failed point and element
    rng4 = np.random.default_rng(SEED + 3)
    pm = fem.structured_hex(3, 2, 2, 3.0, 2.0, 2.0)
    pm = fem.Mesh(pm.xcoord + rng4.uniform(-0.05, 0.05, pm.xcoord.shape), pm.xconn)
    pym = rng4.uniform(1.9e5, 2.0e5, 8 * pm.ne)
    pnu = rng4.uniform(0.28, 0.31, 8 * pm.ne)
    _, _, pNx, pNy, pNz = ref.assemble_m(pm.xcoord, pm.xconn, YM_REF, NU_REF, RHO_REF,
                                         ym=pym, nu=pnu)
    psol = rng4.standard_normal(18 * pm.nn) * 10.0 ** rng4.integers(-6, 2, 18 * pm.nn)
    ptemp = rng4.uniform(-10.0, 50.0, 3 * pm.nn)
    pgp = np.ones(8 * pm.ne, np.int32)
    pgp[3] = 0
    pel = np.ones(pm.ne, np.int32)
    pel[1] = 0
    pD = rng4.uniform(0.0, 0.3, 8 * pm.ne)
    pD[:4] = [0.0, 1e-300, 0.5, 12345.678]
    pws = rng4.uniform(0.0, 1.0, 8 * pm.ne) * 10.0 ** rng4.integers(-300, 300, 8 * pm.ne)
    ptc = np.array([0.5, 0.75, 1.0])
    files = ref.mpipostprocess("job", pm.xconn, pNx, pNy, pNz, pym, pnu, 2 * np.pi * 20, ptc,
                               psol, ptemp, pgp, pel, pD, pws, 2560, 256)
    g.update(pp_xcoord=pm.xcoord, pp_xconn=pm.xconn, pp_ym=pym, pp_nu=pnu, pp_sol=psol,
             pp_temp=ptemp, pp_gpsta=pgp, pp_elsta=pel, pp_D=pD, pp_ws=pws, pp_tc=ptc)
    for k, v in files.items():
        g[f"pp_file_{k}"] = np.array(v)
    np.savez_compressed(os.path.join(HERE, "golden_ref.npz"), **g)
    print("wrote", os.path.join(HERE, "golden_ref.npz"))


if __name__ == "__main__":
    main()
