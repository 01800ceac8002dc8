"""The live path's host-array GMRES, mpipgmres (kronecker_mumps.f90:186-415), through the
C ABI with the exact banded factor in place of MUMPS's, against the oracle's kronpgmres
with the exact LU factor (ilut without dropping): the same algorithm with the same
preconditioner up to rounding.  On this system (cond 6.4e6, 37 iterations) the oracle's
own solution at eps 1e-12 has true relative residual 6.7e-11 and lies 7.5e-12 from a
dense solve, so the bounds below (1e-9 residual, 1e-8 from the oracle) hold with margin."""
import numpy as np
import pytest

from conftest import rel, structured_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def xt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import api
    return api


def test_mpipgmres_band_factor(xt, oracle_lib):
    from xtfem_amd import device, slab
    S = structured_system(8, 4, 3, 16.0, 8.0, 6.0)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * 20.0)     # 10 whole load periods per slab
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    p = S["p"]
    rhs = np.random.default_rng(11).standard_normal(6 * p)
    pr = device.BandPreconditioner(S["Kb"])
    sol, ierr, its, hist = xt.mpipgmres(A, B, S["Kb"], S["Mb"], pr, rhs, im=30, eps=1e-12,
                                        maxits=300)
    assert ierr == 0 and its > 0
    r = rhs - oracle_lib.kronspmv(A, B, S["Kb"], S["Mb"], sol)
    assert np.linalg.norm(r) / np.linalg.norm(rhs) < 1e-9
    lu = oracle_lib.ilut(S["Kb"], p, 0.0)[:3]
    ref, ie, _, _ = oracle_lib.kronpgmres(A, B, S["Kb"], S["Mb"], lu, rhs, im=30, eps=1e-12,
                                          maxits=300)
    assert ie == 0 and rel(sol, ref) < 1e-8
    # history starts at 1 and ends below eps (ro/ro0, kronecker_mumps.f90:270-271, 344)
    assert abs(hist[0] - 1.0) < 1e-14 and hist[-1] <= 1e-12
    # a zero right-hand side: ierr -1, no iteration, zero solution
    z, ierr0, its0, _ = xt.mpipgmres(A, B, S["Kb"], S["Mb"], pr, np.zeros(6 * p))
    assert ierr0 == -1 and its0 == 0 and not np.any(z)
