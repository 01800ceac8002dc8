"""The fatigue slab loop on the device (xtfem_amd.fatigue: slab solve, damage over the
slab, element deletion; main3d.f90:625-691, 803-898) against the oracle.

A notched 10 x 10 x 4 mm block (C3D8, notch 3 x 1 elements at mid-height), clamped at
y = 0, under a 20 Hz tensile traction of 300 MPa on y = 10, with the damage law's
fatigue limit raised to 350 MPa and the critical damage to 0.95 (the fatigue-damage
configuration's material): damage starts after about seven slabs and elements at the
notch tip fail within twelve.  Each slab's solution is the GPU's FP32-IR solve (true
relative residual <= 1e-8); the oracle's damage functor (kokkos_kernel.h:44-344) and
element status (delelement.f90:109-136) are run on the same solutions, so the damage
state must agree bit for bit and the deleted sets must be identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAMS = np.array([350.0, 577.0, 0.08, 1740.0, 0.2, 0.5, 0.5])


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


@pytest.mark.parametrize("readd", [False, True])
def test_fatigue_loop_matches_oracle(dev, oracle_lib, readd):
    from xtfem_amd import fem, slab
    from xtfem_amd.fatigue import FatigueSlabs
    nx, ny, nz = 10, 10, 4
    keep = np.ones((nx, ny, nz), bool)
    keep[0:3, ny // 2, :] = False
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep)
    K, M, Nx, Ny, Nz = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    bc = fem.clamp_mask(mesh, 1, [0.0])
    unit = fem.traction_load(mesh, 1, float(ny), 1, float(nx * nz))
    nip, Dc, dt, f = 256, 0.95, 0.5, 20.0
    fs = FatigueSlabs(mesh, K, M, (Nx, Ny, Nz), bc, unit, 300.0, f, dt, nip, Dc, 2.0e5, 0.3,
                      params=PARAMS, mode=1, readd=readd)
    ne, ngp = mesh.ne, 8 * mesh.ne
    nodes = np.repeat(mesh.xconn, 8, axis=0).ravel()
    st = {k: np.zeros(6 * ngp) for k in ("eps_old", "eps_pl", "sig_eff", "sig_back")}
    st.update({k: np.zeros(ngp) for k in ("p", "D", "ws", "del_po", "sig_d")})
    elesta = np.ones(ne, dtype=np.int32)
    gpsta = np.ones(ngp, dtype=np.int32)
    deleted_oracle = []
    oracle_lib.damage_params(PARAMS)
    try:
        for n in range(12):
            # the oracle's element status from its own damage state, before the slab
            nefail, elesta, gpsta = oracle_lib.chkelesta(ne, 8, Dc, st["D"], elesta, gpsta)
            rec = fs.step()
            assert rec["failed"] == list(nefail), (n, rec["failed"], list(nefail))
            deleted_oracle += list(nefail)
            assert rec["ierr"] == 0 and rec["relres"] <= 1e-8, rec
            sol = fs.sol.cpu().numpy()
            t1 = n * dt
            Nt = slab.damage_nt(np.array([t1, t1 + dt / 2, t1 + dt]), 2 * np.pi * f, nip)
            st = oracle_lib.kokkos_host(ne, mesh.nn, nip, nodes, Nx.ravel(), Ny.ravel(),
                                        Nz.ravel(), Nt.ravel(), sol, np.full(ngp, 2.0e5),
                                        np.full(ngp, 0.3), st)
            st["D"][gpsta == 0] = 0.0            # kokkos_damage.f90:102-107
            st["ws"][gpsta == 0] = 0.0
            got = fs.damage.host_state()
            for k in st:
                assert np.array_equal(got[k], st[k]), (n, k)
    finally:
        oracle_lib.damage_params(None)
    assert len(deleted_oracle) > 0 and np.array_equal(fs.deleted(), deleted_oracle)
