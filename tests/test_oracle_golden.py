"""The oracle (and the product's host-side setup code) against the golden fixtures
written from the compiled reference (tests/golden/make_golden.py).  Runs anywhere:
needs neither the reference nor a GPU."""
import numpy as np
import pytest

from conftest import RHO_REF, YM_REF, NU_REF, digest, rel


def test_assembly_matches_reference_bit_for_bit(rod, golden):
    """fem.assemble restates mpiglbmtx_m + tricsr: identical CSR and Nx/Ny/Nz bits."""
    assert digest(*rod["K"]) == str(golden["asm_K_digest"])
    assert digest(*rod["M"]) == str(golden["asm_M_digest"])
    assert digest(rod["Nx"].reshape(-1, 8), rod["Ny"].reshape(-1, 8),
                  rod["Nz"].reshape(-1, 8)) == str(golden["asm_Nxyz_digest"])
    assert len(rod["K"][2]) == int(golden["asm_K_nnz"])


def test_applybc_matches_reference(rod, golden):
    assert np.array_equal(rod["bc"], golden["bc_mask"])
    assert digest(*rod["Kb"]) == str(golden["Kb_digest"])
    assert digest(*rod["Mb"]) == str(golden["Mb_digest"])


@pytest.mark.parametrize("i", range(4))
def test_stcoeff_m(oracle_lib, golden, i):
    from xtfem_amd import slab
    t1, dt, w = golden["st_pts"][i]
    o = oracle_lib.stcoeff_m(t1, dt, w)
    s = slab.stcoeff_m(t1, dt, w)
    for k in ("k", "m", "ktn", "mtn", "ktn1", "mtn1"):
        g = golden[f"st{i}_{k}"]
        assert rel(o[k], g) < 2e-15, k
        assert rel(s[k], g) < 2e-15, k


def test_stcoeff_t(oracle_lib, golden):
    from xtfem_amd import slab
    o = oracle_lib.stcoeff_t(0.0, 0.5)
    s = slab.stcoeff_t(0.0, 0.5)
    for k in ("k", "m", "ktn", "mtn", "ktn1", "mtn1"):
        assert np.array_equal(o[k], golden[f"stt_{k}"]), k
        assert np.array_equal(s[k], golden[f"stt_{k}"]), k


@pytest.mark.parametrize("i", range(4))
def test_getfextc_m(oracle_lib, golden, i):
    from xtfem_amd import slab
    t1, dt, w = golden["st_pts"][i]
    f, g = oracle_lib.getfextc_m(t1, w, dt)
    fs, gs = slab.getfextc_m(t1, w, dt)
    assert rel(f, golden[f"fx{i}_fextc"]) < 1e-15 and rel(fs, golden[f"fx{i}_fextc"]) < 1e-15
    assert np.array_equal(g, golden[f"fx{i}_fint"]) and np.array_equal(gs, golden[f"fx{i}_fint"])


def test_ilut_matches_reference_bit_for_bit(oracle_lib, rod, golden):
    alu, jlu, ju, ierr = oracle_lib.ilut(rod["Kb"], 20, 1e-4)
    assert ierr == 0 and len(alu) == int(golden["ilut20_nnz"])
    assert digest(alu, jlu, ju) == str(golden["ilut20_digest"])
    alu, jlu, ju, ierr = oracle_lib.ilut(rod["Kb"], 5, 1e-3)
    assert ierr == 0 and digest(alu, jlu, ju) == str(golden["ilut5_digest"])


def test_kron_ops(oracle_lib, rod, golden):
    A, B = golden["A"], golden["B"]
    assert np.array_equal(oracle_lib.kronspmv(A, B, rod["Kb"], rod["Mb"], golden["kron_x"]),
                          golden["kron_y"])
    y = oracle_lib.kronspgemv(A, B, 0.7, -1.3, rod["Kb"], rod["Mb"], golden["kron_x"],
                              golden["gemv_y0"])
    assert np.array_equal(y, golden["gemv_y"])


def test_lusol_kronlusol(oracle_lib, rod, golden):
    alu, jlu, ju, _ = oracle_lib.ilut(rod["Kb"], 20, 1e-4)
    assert np.array_equal(oracle_lib.lusol(alu, jlu, ju, golden["lusol_y"]), golden["lusol_x"])
    # LAPACK's DGETRI and the restated inverse agree to rounding only
    assert rel(oracle_lib.kronlusol(golden["A"], alu, jlu, ju, golden["kron_x"]),
               golden["klu_y"]) < 1e-13


def test_kronpgmres(oracle_lib, rod, golden):
    prec = oracle_lib.ilut(rod["Kb"], 20, 1e-4)[:3]
    A, B = golden["A"], golden["B"]
    for eps, tag in ((1e-8, "e8"), (1e-12, "e12")):
        sol, ierr, its, hist = oracle_lib.kronpgmres(A, B, rod["Kb"], rod["Mb"], prec,
                                                     golden["gmres_rhs"], im=30, eps=eps,
                                                     maxits=3000)
        assert ierr == 0
        ref = golden[f"gmres_{tag}_sol"]
        # same algorithm; only BLAS/LAPACK rounding differs, amplified by the solve
        assert rel(sol, ref) < (1e-6 if tag == "e8" else 1e-9), tag


def test_damage_step_vs_getdfinal(oracle_lib, golden):
    """The oracle's return map (restating kokkos_kernel.h:85-341) against the
    reference's Fortran twin getdfinal.f, one step at a time from the same state
    (whole trajectories are not comparable: the two sum the equivalent stress in a
    different order and the return map is discontinuous at load reversals)."""
    n = len(golden["dmg_strain"])
    plastic = damaged = 0
    for i in range(n):
        eo, ep, se, sb = golden["dmg_in6"][i]
        sc = golden["dmg_in1"][i]
        o = oracle_lib.damage_step(golden["dmg_strain"][i], YM_REF, NU_REF, eo, ep, se, sb, sc)
        r6, r1 = golden["dmg_out6"][i], golden["dmg_out1"][i]
        for j in range(4):
            assert rel(o[j], r6[j]) < 1e-11, (i, j)
        assert np.all(np.abs(o[4] - r1) <= 1e-11 * np.maximum(np.abs(r1), 1e-30)), i
        plastic += r1[0] != sc[0]
        damaged += r1[1] > sc[1]
    assert plastic > 100 and damaged > 20


def test_chkelesta(oracle_lib, golden, rod):
    nefail, elesta, gpsta = oracle_lib.chkelesta(rod["mesh"].ne, 8, 0.3, golden["chk_D"])
    assert np.array_equal(nefail, golden["chk_nefail"])
    assert np.array_equal(elesta, golden["chk_elesta"])
    assert np.array_equal(gpsta, golden["chk_gpsta"])
