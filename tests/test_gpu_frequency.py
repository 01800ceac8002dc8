"""Slab solves across load frequencies (per-case temporal matrices A_r, B_r for
f_r in 10-100 Hz) with the shifted-factor split preconditioner, against the oracle's
FP64 kronpgmres (kronecker.f90:181-365, exact K^-1 preconditioner) converged to 1e-12.

The dominant eigenvalue of A_r^-1 B_r grows like f_r^4 (5.4e4 at 10 Hz, 5.3e8 at
100 Hz for the half-second slabs used here).  Once it reaches the structure's lowest
K^-1 M eigenvalues, A_r^-1 (x) K^-1 ignores a mass term of the size of K and GMRES with a
factor rounded to FP32 stagnates.  The test plate is a notched 40 x 20 x 4 mm plate of a
material with steel's stiffness and 50 times its density, so that its lowest mode sits
well below the 100 Hz cases' dominant temporal eigenvalue (the 100 x 46 x 4 mm steel
plate's does, at about a sixth of it).  Bars: every case converges to true relative
residual 1e-8, and its displacement lies within 1e-6 (relative, max norm) of the
oracle's."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

RHO_HEAVY = 50 * 7.85e-9


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


@pytest.fixture(scope="module")
def sweep(oracle_lib):
    from xtfem_amd import fem, slab
    nx, ny, nz, notch = 40, 20, 4, 8
    keep = np.ones((nx, ny, nz), bool)
    keep[0:notch, ny // 2, :] = False
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep)
    K, M, *_ = fem.assemble(mesh, 2.0e5, 0.3, RHO_HEAVY)
    bc = fem.clamp_mask(mesh, 1, [0.0])
    Kb, Mb = fem.applybc(K, bc), fem.applybc(M, bc)
    p = 3 * mesh.nn
    R = 8
    rng = np.random.default_rng(20261205)
    freq = np.linspace(13.0, 97.0, R) + rng.uniform(-2.0, 2.0, R)   # distinct, 10-100 Hz
    amp = rng.uniform(20.0, 120.0, R)
    w = 2 * np.pi * freq
    dt = np.round(0.5 * freq) / freq
    c = slab.stcoeff_m(0.0, dt, w)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    unit = fem.traction_load(mesh, 1, float(ny), 1, float(nx * nz))
    fx, _ = slab.getfextc_m(0.0, w, dt)
    rhs = np.array([slab.slab_rhs(p, fx[r] * amp[r], None, unit, fixed=np.tile(bc, 6))
                    for r in range(R)])
    bf = oracle_lib.BandFactor(Kb)
    xref = []
    for r in range(R):
        x, ierr, its, _ = oracle_lib.kronpgmres(A[r], B[r], Kb, Mb, bf, rhs[r], im=30,
                                                eps=1e-12, maxits=3000)
        assert ierr == 0, (r, its)
        xref.append(x)
    return dict(Kb=Kb, Mb=Mb, p=p, A=A, B=B, rhs=rhs, xref=np.array(xref), freq=freq, bc=bc)


def test_shift_choice(dev, sweep):
    """The plate's lowest mode (device power iteration on K^-1 M) and the shifts the
    cases need: the high-frequency cases' dominant temporal eigenvalues exceed it."""
    pr0 = dev.BandPreconditioner(sweep["Kb"])
    mu = dev.lowest_mode(pr0, sweep["Mb"], sweep["bc"])
    lam = dev.dominant_temporal_mode(sweep["A"], sweep["B"])
    assert np.all(np.isfinite(lam))
    assert 3e7 < mu < 8e7, mu                      # 2.47e9 / 50 for this plate and density
    assert lam.max() > 4 * mu
    sh = dev.shifts_for_cases(lam, mu)
    assert 1 <= len(sh) <= 4


@pytest.mark.parametrize("mode,eps,xtol,fp32_only", [(1, 1e-8, 1e-7, False), (1, 1e-8, 1e-7, True),
                                                     (0, 1e-10, 0.0, False),
                                                     (2, 1e-10, 0.0, False)])
def test_frequency_sweep_matches_oracle(dev, sweep, mode, eps, xtol, fp32_only):
    """FP32-IR (forward-error test 1e-7; shifted factors in FP64 and FP32, or FP32 only),
    FP64 GMRES and FP64 FGMRES with FP32 sweeps, all with the split preconditioner: every
    case converges and matches the oracle."""
    pr0 = dev.BandPreconditioner(sweep["Kb"])
    mu = dev.lowest_mode(pr0, sweep["Mb"], sweep["bc"])
    lam = dev.dominant_temporal_mode(sweep["A"], sweep["B"])
    pr = dev.BandPreconditioner(sweep["Kb"], M=sweep["Mb"],
                                shifts=dev.shifts_for_cases(lam, mu), shifted_fp32_only=fp32_only)
    assert len(pr.shifts) >= 1
    op = dev.KronOperator([sweep["Kb"], sweep["Mb"]], q=6)
    x, its, ierr, relres = dev.slab_solve(op, pr, sweep["A"], sweep["B"], sweep["rhs"], im=30,
                                          eps=eps, maxits=300, mode=mode, xtol=xtol)
    for r in range(len(x)):
        err = rel(x[r], sweep["xref"][r])
        assert ierr[r] == 0 and relres[r] <= max(eps, 1e-8), (r, sweep["freq"][r], its[r], relres[r])
        assert err <= 1e-6, (r, sweep["freq"][r], its[r], err)
        assert its[r] <= 40, (r, sweep["freq"][r], its[r])


def test_split_fp32_converges_where_unsplit_stagnates(dev, sweep):
    """The failure the split removes: FP32-IR with A^-1 (x) K^-1 alone needs many more
    iterations for the highest-frequency case than with the split (same tolerances)."""
    r = [len(sweep["freq"]) - 1]
    op = dev.KronOperator([sweep["Kb"], sweep["Mb"]], q=6)
    pr0 = dev.BandPreconditioner(sweep["Kb"])
    mu = dev.lowest_mode(pr0, sweep["Mb"], sweep["bc"])
    lam = dev.dominant_temporal_mode(sweep["A"], sweep["B"])
    pr = dev.BandPreconditioner(sweep["Kb"], M=sweep["Mb"], shifts=dev.shifts_for_cases(lam, mu))
    args = (sweep["A"][r], sweep["B"][r], sweep["rhs"][r])
    _, its_u, _, _ = dev.slab_solve(op, pr0, *args, eps=1e-8, maxits=300, mode=2)
    x, its_s, ierr, _ = dev.slab_solve(op, pr, *args, eps=1e-8, maxits=300, mode=2)
    assert ierr[0] == 0 and its_s[0] <= 15 and its_u[0] >= 2 * its_s[0], (its_u, its_s)
