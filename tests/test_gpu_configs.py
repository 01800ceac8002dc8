"""GPU parity at the benchmark configurations' full sizes.

* single-edge-notched plate 100 x 46 x 4 (p = 71,205 before the notch): batched FP32-IR
  slab solves across 10-100 Hz with the split preconditioner, against the oracle's FP64
  kronpgmres (kronecker.f90:181-365, exact K^-1 through LAPACK's band Cholesky) at 1e-12;
* stiffened panel 240 x 200 x 2 (p = 435,969): the Kron-SpMM at 32 load cases against
  the oracle's kronspmv (kronecker.f90:19-85) on sampled cases, and two consecutive
  batched slab solves whose true FP64 residual the oracle recomputes;
* fatigue block 60 x 60 x 12 (345,600 Gauss points): the damage kernel with the 350 MPa
  material against the oracle's functor (kokkos_kernel.h:44-344) bit for bit on a
  strided sample of elements, and the element status against brute force.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


def test_plate_frequency_sweep_vs_oracle(dev, oracle_lib):
    from xtfem_amd import fem, slab
    nx, ny, nz, notch = 100, 46, 4, 20
    keep = np.ones((nx, ny, nz), bool)
    keep[0:notch, ny // 2, :] = False
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep)
    K, M, *_ = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    bc = fem.clamp_mask(mesh, 1, [0.0])
    Kb, Mb = fem.applybc(K, bc), fem.applybc(M, bc)
    p = 3 * mesh.nn
    freq = np.array([12.0, 37.0, 63.0, 96.0])
    amp = np.array([80.0, 40.0, 110.0, 25.0])
    dt = np.round(0.5 * freq) / freq
    w = 2 * np.pi * freq
    c = slab.stcoeff_m(0.0, dt, w)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    unit = fem.traction_load(mesh, 1, float(ny), 1, float(nx * nz))
    fx, _ = slab.getfextc_m(0.0, w, dt)
    rhs = np.array([slab.slab_rhs(p, fx[r] * amp[r], None, unit, fixed=np.tile(bc, 6))
                    for r in range(4)])
    op = dev.KronOperator([Kb, Mb], q=6)
    pr0 = dev.BandPreconditioner(Kb)
    mu = dev.lowest_mode(pr0, Mb, bc)
    assert 8e7 < mu < 1.1e8, mu                # the plate's first mode, 1537 Hz
    del pr0
    lam = dev.dominant_temporal_mode(A, B)
    pr = dev.BandPreconditioner(Kb, M=Mb, shifts=dev.shifts_for_cases(lam, mu))
    x, its, ierr, relres = dev.slab_solve(op, pr, A, B, rhs, eps=1e-8, maxits=300, mode=1,
                                          xtol=1e-7)
    lf = oracle_lib.LapackBandFactor(Kb)
    for r in (0, 3):                           # the lowest and the highest frequency
        xo, ie, _, _ = oracle_lib.kronpgmres(A[r], B[r], Kb, Mb, lf, rhs[r], eps=1e-12,
                                             maxits=500)
        assert ie == 0
        assert ierr[r] == 0 and relres[r] <= 1e-8, (r, its[r], relres[r])
        assert rel(x[r], xo) <= 1e-6, (r, rel(x[r], xo))
    assert np.all(ierr == 0) and np.all(relres <= 1e-8)


@pytest.fixture(scope="module")
def panel(dev):
    from xtfem_amd import device, fem
    nx, ny, nz = 240, 200, 2
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz))
    x = mesh.xcoord
    clamp = (x[:, 0] == 0.0) | (x[:, 0] == float(nx)) | (x[:, 1] == 0.0)
    bc = np.repeat(clamp, 3).astype(np.int32)
    asm = device.Assembly(mesh, bc)
    asm.run(2.0e5, 0.3, 7.85e-9)
    Kb, Mb = asm.host()
    unit = fem.traction_load(mesh, 1, float(ny), 1, float(nx * nz))
    return dict(mesh=mesh, bc=bc, Kb=Kb, Mb=Mb, p=3 * mesh.nn, unit=unit)


def test_panel_spmm_32_cases_vs_oracle(dev, oracle_lib, panel):
    from xtfem_amd import slab
    R, p = 32, panel["p"]
    assert p == 435969
    rng = np.random.default_rng(20261206)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * rng.uniform(10, 100, R))
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    X = rng.standard_normal((R, 6 * p))
    op = dev.KronOperator([panel["Kb"], panel["Mb"]], q=6)
    assert op.format == "blocks"
    for fp32, tol in ((False, 1e-12), (True, 2e-5)):
        Y = op.apply_host(A, B, X, fp32=fp32)
        for r in (0, 13, 31):
            ref = oracle_lib.kronspmv(A[r], B[r], panel["Kb"], panel["Mb"], X[r])
            assert rel(Y[r], ref) < tol, (fp32, r, rel(Y[r], ref))


def test_panel_two_batched_slabs_true_residual(dev, oracle_lib, panel):
    """Slab 0 and slab 1 (jump term from slab 0) of 8 load cases: every case's true
    residual, recomputed by the oracle, is <= 1e-8 of its right-hand side."""
    import torch
    from xtfem_amd import slab
    R, p, bc = 8, panel["p"], panel["bc"]
    rng = np.random.default_rng(20261207)
    freq, amp = rng.uniform(10, 100, R), rng.uniform(20, 120, R)
    dt = np.round(0.5 * freq) / freq
    w = 2 * np.pi * freq
    c = slab.stcoeff_m(0.0, dt, w)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    op = dev.KronOperator([panel["Kb"], panel["Mb"]], q=6)
    pr0 = dev.BandPreconditioner(panel["Kb"])
    mu = dev.lowest_mode(pr0, panel["Mb"], bc)
    del pr0
    torch.cuda.synchronize()
    pr = dev.BandPreconditioner(panel["Kb"], M=panel["Mb"],
                                shifts=dev.shifts_for_cases(dev.dominant_temporal_mode(A, B),
                                                            mu, 16.0),
                                shifted_fp32_only=True)
    xprev = np.zeros((R, 6 * p))
    for n in range(2):
        fx, _ = slab.getfextc_m(n * dt, w, dt)
        rhs = np.array([slab.slab_rhs(p, fx[r] * amp[r], None, panel["unit"],
                                      jump=oracle_lib.kronspmv(c["ktn1"][r], c["mtn1"][r],
                                                               panel["Kb"], panel["Mb"],
                                                               xprev[r]),
                                      fixed=np.tile(bc, 6)) for r in range(R)])
        x, its, ierr, relres = dev.slab_solve(op, pr, A, B, rhs, eps=1e-8, maxits=300,
                                              mode=1, xtol=1e-7)
        for r in range(R):
            res = rhs[r] - oracle_lib.kronspmv(A[r], B[r], panel["Kb"], panel["Mb"], x[r])
            assert ierr[r] == 0
            assert np.linalg.norm(res) <= 1.01e-8 * np.linalg.norm(rhs[r]), (n, r, its[r])
        xprev = x


def test_fatigue_block_damage_kernel_vs_oracle(dev, oracle_lib):
    import torch
    from xtfem_amd import fem, slab
    nx, ny, nz = 60, 60, 12
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz))
    _, _, Nx, Ny, Nz = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    params = np.array([350.0, 577.0, 0.08, 1740.0, 0.2, 0.5, 0.5])
    # seeded field: stretch and bend, peak von Mises about 1.2 x 350 MPa at the corner
    rng = np.random.default_rng(20261208)
    xi = mesh.xcoord / np.array([nx, ny, nz], float)
    wk = rng.uniform(0.5, 1.0, 6) * np.array([1, 1, 1, 0.6, 0.6, 0.6])
    eps0 = 2.0e-3
    sol = np.zeros((6, mesh.nn, 3))
    for k in range(6):
        sol[k, :, 0] = eps0 * wk[k] * nx * xi[:, 0] * (1 + 0.3 * xi[:, 1])
        sol[k, :, 1] = -0.3 * eps0 * wk[k] * ny * xi[:, 1] + 0.2 * eps0 * wk[k] * nx * xi[:, 0] ** 2
        sol[k, :, 2] = -0.3 * eps0 * wk[k] * nz * xi[:, 2]
    nip = 256
    Nt = slab.damage_nt(np.array([0.0, 0.25, 0.5]), 2 * np.pi * 20.0, nip)
    ds = dev.DamageState(mesh, Nx, Ny, Nz, 2.0e5, 0.3, params=params)
    nslab = 3
    for _ in range(nslab):
        ds.step(torch.as_tensor(sol.ravel(), device="cuda"),
                torch.as_tensor(Nt.ravel(), device="cuda"), nip)
    got = ds.host_state()
    assert np.mean(got["p"] > 0) > 0.01          # the sample holds yielding points
    sample = np.arange(0, mesh.ne, 97)
    ngs = 8 * len(sample)
    gp = (8 * sample[:, None] + np.arange(8)).ravel()
    st0 = {k: np.zeros(6 * ngs) for k in ("eps_old", "eps_pl", "sig_eff", "sig_back")}
    st0.update({k: np.zeros(ngs) for k in ("p", "D", "ws", "del_po", "sig_d")})
    oracle_lib.damage_params(params)
    try:
        st = st0
        for _ in range(nslab):
            st = oracle_lib.kokkos_host(len(sample), mesh.nn, nip,
                                        np.repeat(mesh.xconn[sample], 8, axis=0).ravel(),
                                        Nx[gp].ravel(), Ny[gp].ravel(), Nz[gp].ravel(),
                                        Nt.ravel(), sol.ravel(), np.full(ngs, 2.0e5),
                                        np.full(ngs, 0.3), st)
    finally:
        oracle_lib.damage_params(None)
    for k in st:
        w6 = 6 if k in ("eps_old", "eps_pl", "sig_eff", "sig_back") else 1
        g = got[k].reshape(-1, w6)[gp].ravel()
        assert np.array_equal(g, st[k]), k
    # element status on the full block against brute force, at a Dc the field reaches
    Dc = float(np.quantile(got["D"][got["D"] > 0], 0.5)) if np.any(got["D"] > 0) else 1e-30
    failed = ds.failed_elements(Dc)
    brute = np.flatnonzero((got["D"].reshape(mesh.ne, 8) >= Dc).any(axis=1)) + 1
    assert np.array_equal(failed, brute)
