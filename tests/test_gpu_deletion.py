"""GPU side of element deletion (main3d.f90:625-691): the device element-status
reduction (chkelesta, delelement.f90:109-136) against the reference's golden
vectors, the zeroing of failed Gauss points (kokkos_damage.f90:102-107), and slab
solves on the device after a deletion event (operator and banded factor rebuilt from
the edited matrices) against the oracle's FP64 GMRES with the exact factor."""
import numpy as np
import pytest

from conftest import NU_REF, YM_REF, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


def test_failed_elements_on_device_golden(dev, rod, golden):
    import torch
    ds = dev.DamageState(rod["mesh"], rod["Nx"], rod["Ny"], rod["Nz"], YM_REF, NU_REF)
    ds.s["D"].copy_(torch.as_tensor(golden["chk_D"], device="cuda"))
    nef = ds.failed_elements(0.3)
    assert np.array_equal(nef, golden["chk_nefail"])
    assert np.array_equal(ds.elesta.cpu().numpy(), golden["chk_elesta"])
    assert np.array_equal(ds.gpsta.cpu().numpy(), golden["chk_gpsta"])
    # failed Gauss points lose D and ws; the others keep them
    ws = torch.rand(8 * rod["mesh"].ne, dtype=torch.float64, device="cuda")
    ds.s["ws"].copy_(ws)
    from xtfem_amd._lib import check, lib
    check(lib().xt_damage_zero_failed(8 * rod["mesh"].ne, ds.gpsta.data_ptr(),
                                      ds.s["D"].data_ptr(), ds.s["ws"].data_ptr(), None), "zero")
    torch.cuda.synchronize()
    dead = golden["chk_gpsta"] == 0
    D = ds.s["D"].cpu().numpy()
    W = ds.s["ws"].cpu().numpy()
    assert np.all(D[dead] == 0.0) and np.all(W[dead] == 0.0)
    assert np.array_equal(D[~dead], golden["chk_D"][~dead])
    assert np.array_equal(W[~dead], ws.cpu().numpy()[~dead])
    # a second pass finds nothing new below Dc
    assert ds.failed_elements(0.3).size == 0


@pytest.mark.parametrize("readd", [True, False])
def test_slab_solve_after_deletion(dev, oracle_lib, readd):
    """Notched plate clamped at y = 0: a corner element on the free edge (its corner node
    becomes an orphan) and two elements at the notch tip fail; the operator and banded
    factor are rebuilt from the event's matrices and the batched slab solve meets the
    true FP64 residual 1e-9 and 1e-6 relative displacement against the oracle."""
    import torch
    from xtfem_amd import fem, slab
    from xtfem_amd.deletion import ElementDeletion
    nx, ny, nz, notch = 14, 8, 2, 4
    keep = np.ones((nx, ny, nz), bool)
    keep[0:notch, ny // 2, :] = False
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep)
    K, M, Nx, Ny, Nz = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    bc = fem.clamp_mask(mesh, 1, [0.0])
    p = 3 * mesh.nn
    deg = fem.node_degree(mesh)
    ytop = mesh.xcoord[:, 1].max()
    corner = [n for n in np.flatnonzero(deg == 1) if mesh.xcoord[n, 1] == ytop][0] + 1
    e_corner = int(np.flatnonzero((mesh.xconn == corner).any(axis=1))[0]) + 1
    cen = mesh.xcoord[mesh.xconn - 1].mean(axis=1)
    tip = np.flatnonzero((np.abs(cen[:, 0] - (notch + 0.5)) < 0.01) &
                         (np.abs(cen[:, 1] - (ny // 2 + 0.5)) < 0.6))[:2] + 1
    failed = np.unique(np.r_[e_corner, tip]).astype(np.int32)

    ds = dev.DamageState(mesh, Nx, Ny, Nz, 2.0e5, 0.3)
    D = np.zeros(8 * mesh.ne)
    D[8 * (failed - 1) + 3] = 0.31
    ds.s["D"].copy_(torch.as_tensor(D, device="cuda"))
    dl = ElementDeletion(mesh, K, M, bc, 2.0e5, 0.3, Dc=0.3, readd=readd)
    ev = dl.step(ds)
    assert ev is not None and np.array_equal(ev.nefail, failed)
    assert ev.ndfail.tolist() == [corner]
    assert np.all(ev.fixed[3 * (corner - 1):3 * corner] == 1)

    R = 3
    rng = np.random.default_rng(11)
    amps, freqs = rng.uniform(20, 120, R), rng.uniform(10, 60, R)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * freqs)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    rext0 = fem.traction_load(mesh, 1, ytop, 1, 1.0)
    fx, _ = slab.getfextc_m(0.0, 2 * np.pi * freqs, 0.5)
    fixed6 = np.tile(ev.fixed, 6)
    rhs = np.array([slab.slab_rhs(p, fx[r], None, amps[r] * rext0, fixed=fixed6)
                    for r in range(R)])
    op = dev.KronOperator([ev.K, ev.M], q=6)
    pr = dev.BandPreconditioner(ev.K)
    for mode in (0, 1):
        x, its, ierr, relres = dev.slab_solve(op, pr, A, B, rhs, im=30, eps=1e-10, maxits=200,
                                              mode=mode)
        assert np.all(ierr == 0), (mode, ierr, its)
        p_ = len(ev.K[0]) - 1
        alu, jlu, ju, ie = oracle_lib.ilut(ev.K, p_, 0.0)
        assert ie == 0
        for r in range(R):
            y = oracle_lib.kronspmv(A[r], B[r], ev.K, ev.M, x[r])
            rr = np.linalg.norm(rhs[r] - y) / np.linalg.norm(rhs[r])
            assert rr < 1e-9, (mode, r, rr)
            ref, ie, _, _ = oracle_lib.kronpgmres(A[r], B[r], ev.K, ev.M, (alu, jlu, ju), rhs[r],
                                                  im=30, eps=1e-12, maxits=100)
            assert ie == 0 and rel(x[r], ref) < 1e-6
            assert np.all(np.abs(x[r][fixed6 != 0]) <= 1e-12 * np.abs(x[r]).max())
