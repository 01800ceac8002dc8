"""Thermal side on the GPU: conductivity/capacity assembly (xt_asm_thermal), the
per-slab material update and thermal load (xt_asm_thermo_load) against the compiled
reference's vectors (tests/golden) and the host restatements, bit for bit; and the
coupled thermo-mechanical slab loop (xtfem_amd.thermo) against the oracle's loop with
direct solves (temperature to 1e-8, displacement to 1e-6 relative)."""
import numpy as np
import pytest

from conftest import NU_REF, RHO_REF, YM_REF, digest, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


def compact(ia, ja, a):
    """Drop stored zeros (tricsr/csrapplybc drop them; the device pattern keeps them)."""
    ia, ja, a = (np.asarray(x) for x in (ia, ja, a))
    n = len(ia) - 1
    row = np.repeat(np.arange(n), np.diff(ia))
    keep = a != 0.0
    cnt = np.bincount(row[keep], minlength=n)
    return (np.r_[0, np.cumsum(cnt)].astype(np.int32) + 1, ja[keep].astype(np.int32),
            a[keep].astype(np.float64))


def test_thermal_matrices(dev, rod, golden):
    from xtfem_amd import fem
    mesh = rod["mesh"]
    asm = dev.Assembly(mesh, np.zeros(rod["p"], dtype=np.int32))
    tf = np.zeros(mesh.nn, dtype=np.int32)
    asm.thermal(tf, 0.1, 452.0, RHO_REF)
    ia, ja = asm.T_pattern
    assert digest(*compact(ia, ja, asm.Kt.cpu().numpy())) == str(golden["th_Kt_digest"])
    assert digest(*compact(ia, ja, asm.Mt.cpu().numpy())) == str(golden["th_Mt_digest"])
    # distorted mesh with fixed temperatures: with and without the 1-0 method
    m2 = fem.structured_hex(5, 3, 2, 5.0, 3.0, 2.0)
    rng = np.random.default_rng(4)
    m2 = fem.Mesh(m2.xcoord + rng.uniform(-0.1, 0.1, m2.xcoord.shape), m2.xconn)
    tf2 = fem.fixed_temperature_mask(m2, 0, [m2.xcoord[:, 0].min()], tol=0.2)
    assert tf2.sum() > 0
    a2 = dev.Assembly(m2, np.zeros(3 * m2.nn, dtype=np.int32))
    a2.thermal(tf2, 0.1, 452.0, RHO_REF)
    Kh, Mh = fem.assemble_t(m2, 0.1, 452.0, RHO_REF)
    ia, ja = a2.T_pattern
    for dvals, host in ((a2.Kt0, Kh), (a2.Mt0, Mh), (a2.Kt, fem.applybc(Kh, tf2)),
                        (a2.Mt, fem.applybc(Mh, tf2))):
        for x, y in zip(compact(ia, ja, dvals.cpu().numpy()), host):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("tag,Tref", [("", 0.0), ("r", 4.0)])
def test_thermo_load(dev, rod, golden, tag, Tref):
    import torch
    from xtfem_amd import fem
    mesh = rod["mesh"]
    asm = dev.Assembly(mesh, np.zeros(rod["p"], dtype=np.int32))
    T = golden["th_T"]
    # interleaved thermal solution [nn][3] (one load case): node n, temporal node j
    Ti = torch.as_tensor(T.reshape(3, mesh.nn).T.copy().ravel(), device="cuda")
    ym, nu, vv = asm.thermo_load(Ti, 3, 2, Tref, fem.ALPHA_REF, YM_REF, NU_REF)
    assert np.array_equal(ym.cpu().numpy(), golden[f"th{tag}_ym"])
    assert np.array_equal(nu.cpu().numpy(), golden[f"th{tag}_nu"])
    assert np.array_equal(vv.cpu().numpy(), golden[f"th{tag}_vv"])
    K, M = asm.run(ym, nu, RHO_REF)
    Kh, Mh = asm.host()
    assert digest(*compact(*Kh)) == str(golden[f"th{tag}_K_digest"])
    assert digest(*compact(*Mh)) == str(golden[f"th{tag}_M_digest"])


def coupled_case():
    from xtfem_amd import fem
    mesh = fem.structured_hex(6, 4, 3, 6.0, 4.0, 3.0, origin=(0.0, -2.0, 0.0))
    fixed_m = fem.clamp_mask(mesh, 2, [0.0])
    tfixed = fem.fixed_temperature_mask(mesh, 2, [0.0, 3.0])
    rext_t = fem.heat_flux_load(mesh, 1, 2.0, [2.0, -2.0], 20.0)
    rext_m = fem.traction_load(mesh, 2, 3.0, 2, 400.0) + fem.traction_load(mesh, 2, 3.0, 0, 150.0)
    return mesh, fixed_m, tfixed, rext_m, rext_t


@pytest.mark.parametrize("mode,shift", [(0, 0.0), (0, "auto"), (1, "auto")])
def test_coupled_slabs_match_oracle(dev, oracle_lib, mode, shift):
    from xtfem_amd import fem, thermo
    mesh, fixed_m, tfixed, rext_m, rext_t = coupled_case()
    dt, w, nslab = 0.5, 2 * np.pi * 20.0, 3
    ref = oracle_lib.thermo_slabs(mesh, fixed_m, tfixed, rext_m, rext_t, nslab, dt, w,
                                  thermo.YM0, thermo.NU0, thermo.RHO, thermo.KAPPA, thermo.CP,
                                  fem.ALPHA_REF, 0.0, 0.0)
    run = thermo.ThermoMechanicalSlabs(mesh, fixed_m, tfixed, rext_m, rext_t, dt, w, mode=mode,
                                       thermal_shift=shift)
    for s in range(nslab):
        its_t, its_m = run.step()
        T_ref, u_ref, ym_ref, nu_ref, vv_ref = ref[s]
        T = run.temperature()
        u = run.displacement()
        print(f"slab {s + 1}: its {its_t}/{its_m}, T max {T.max():.3f}, "
              f"rel T {rel(T, T_ref):.2e}, rel u {rel(u, u_ref):.2e}", flush=True)
        assert T.max() > 10.0 and np.abs(u).max() > 0.0
        assert rel(T, T_ref) < 1e-8
        assert rel(u, u_ref) < 1e-6
        assert rel(run.asm.ym_t.cpu().numpy(), ym_ref) < 1e-10
        assert rel(run.asm.vv.cpu().numpy(), vv_ref) < 1e-8
    # the material did change with the temperature
    assert np.ptp(run.asm.ym_t.cpu().numpy()) > 1.0


def test_coupled_result_files(dev, oracle_lib, tmp_path):
    """Result files of a coupled slab: the same records as the host restatement writes
    from the oracle's loop (numbers to the solver tolerance)."""
    from xtfem_amd import fem, postprocess as pp, thermo
    mesh, fixed_m, tfixed, rext_m, rext_t = coupled_case()
    dt, w = 0.5, 2 * np.pi * 20.0
    ref = oracle_lib.thermo_slabs(mesh, fixed_m, tfixed, rext_m, rext_t, 2, dt, w,
                                  thermo.YM0, thermo.NU0, thermo.RHO, thermo.KAPPA, thermo.CP,
                                  fem.ALPHA_REF, 0.0, 0.0)
    run = thermo.ThermoMechanicalSlabs(mesh, fixed_m, tfixed, rext_m, rext_t, dt, w)
    run.step()
    run.step()
    (tmp_path / "d").mkdir()
    (tmp_path / "h").mkdir()
    t = run.write_results(pp.ResultFiles("job", str(tmp_path / "d")), 2560, 256)
    T_ref, u_ref, ym_ref, nu_ref, _ = ref[1]
    _, _, Nx, Ny, Nz = fem.assemble(mesh, ym_ref, nu_ref, thermo.RHO)
    ne = mesh.ne
    t2 = pp.ResultFiles("job", str(tmp_path / "h")).write_step(
        mesh, Nx, Ny, Nz, ym_ref, nu_ref, w, [0.5, 0.75, 1.0], u_ref, T_ref,
        np.ones(8 * ne, np.int32), np.ones(ne, np.int32), np.zeros(8 * ne), np.zeros(8 * ne),
        2560, 256)
    assert t == t2
    for suf in ("time", "disp", "temp", "strain", "stress", "damage", "energy", "status"):
        name = f"job.{suf}" if suf in ("time", "disp", "temp") else f"job_0.{suf}"
        a = (tmp_path / "d" / name).read_text().split()
        b = (tmp_path / "h" / name).read_text().split()
        assert len(a) == len(b), suf
        fa, fb = np.array([float(x) for x in a]), np.array([float(x) for x in b])
        assert rel(fa, fb) < 1e-6, suf


def test_driver_panel_mesh(dev, oracle_lib):
    """The coupled loop on the reference's own panel mesh (inp/panel_7634_3749.inp, kept
    in tests/golden) with the driver's boundary conditions and loads, renumbered by
    reverse Cuthill-McKee for the band factor: temperatures against the oracle's thermal
    slabs with direct solves, mechanical slabs to the solver tolerance, constraints
    held."""
    import os
    from conftest import GOLDEN
    from xtfem_amd import fem, thermo
    with np.load(os.path.join(GOLDEN, "panel_7634_3749.npz"), allow_pickle=False) as z:
        mesh = fem.Mesh(z["xcoord"], z["xconn"])
    mesh = fem.renumber(mesh, fem.rcm_order(mesh))
    fixed_m, tfixed, rext_m, rext_t = thermo.driver_conditions(mesh)
    assert tfixed.sum() > 0 and fixed_m.sum() > 0 and rext_m.sum() > 0 and rext_t.sum() > 0
    dt, w = 0.5, 2 * np.pi * 20.0
    run = thermo.ThermoMechanicalSlabs(mesh, fixed_m, tfixed, rext_m, rext_t, dt, w)
    assert run.pr_m.bandwidth <= 3 * fem.node_bandwidth(mesh) + 2
    ref = oracle_lib.thermo_slabs(mesh, fixed_m, tfixed, rext_m, rext_t, 2, dt, w, thermo.YM0,
                                  thermo.NU0, thermo.RHO, thermo.KAPPA, thermo.CP,
                                  fem.ALPHA_REF, 0.0, 0.0, mechanical=False)
    for s in range(2):
        its_t, its_m = run.step()
        T = run.temperature()
        print(f"panel slab {s + 1}: its {its_t}/{its_m}, T {T.min():.3f}..{T.max():.3f}, "
              f"rel T {rel(T, ref[s][0]):.2e}, relres m {run.last['relres_m']:.2e}", flush=True)
        assert rel(T, ref[s][0]) < 1e-8
        assert run.last["relres_m"] <= 1e-10 and run.last["relres_t"] <= 1e-10
        u = run.displacement().reshape(6, mesh.nn, 3)
        assert np.all(u[:, fixed_m.reshape(-1, 3)[:, 0] != 0, :] == 0.0)
        assert np.abs(u).max() > 0.0
        Tn = T.reshape(3, mesh.nn)
        assert np.all(Tn[:, tfixed != 0] == 0.0) and Tn.max() > 0.0
