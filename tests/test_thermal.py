"""Thermal side of the coupled slab (host restatements in xtfem_amd.fem / slab) against
the compiled reference's vectors (tests/golden/make_golden.py): conductivity and
capacity matrices (mpiglbmtx_t, elemtx_t), the temperature-dependent material
(material_update), the thermal load vv of mpiglbmtx_m (calc_thermo_stress + elevec_m)
and the heat-flux coefficients (getfextc_t); plus checks the mathematics fixes
(partition of unity, conservation, uniform free thermal expansion, the temporal
projection getfextc_t integrates)."""
import numpy as np
import pytest

from conftest import RHO_REF, YM_REF, NU_REF, digest, rel

KAPPA, CP = 0.1, 452.0


def test_shape_tables_t(golden):
    from xtfem_amd import fem
    from oracle import ref
    w, ns, dns = fem.shape_tables_t()
    assert np.allclose(ns.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    assert np.allclose(dns.sum(axis=1), 0.0, rtol=0, atol=1e-15)
    if ref.available():
        for a, b in zip(ref.shape_tables_t(), (w, ns, dns)):
            assert np.array_equal(a, b)


def test_thermal_matrices_bit_for_bit(rod, golden):
    from xtfem_amd import fem
    Kt, Mt = fem.assemble_t(rod["mesh"], KAPPA, CP, RHO_REF)
    assert digest(*Kt) == str(golden["th_Kt_digest"])
    assert digest(*Mt) == str(golden["th_Mt_digest"])
    assert len(Kt[2]) == int(golden["th_Kt_nnz"])


def test_thermal_matrices_physics():
    """A constant temperature carries no flux (K_t 1 = 0); total capacity is
    rho cp volume; K_t is symmetric positive semi-definite with conductivity kappa:
    the linear field T = x gives the energy kappa * volume."""
    import scipy.sparse as sp
    from xtfem_amd import fem
    mesh = fem.structured_hex(4, 3, 2, 4.0, 1.5, 1.0)
    rng = np.random.default_rng(2)
    Kt, Mt = fem.assemble_t(mesh, KAPPA, CP, RHO_REF)
    n = mesh.nn
    K = sp.csr_matrix((Kt[2], Kt[1] - 1, Kt[0] - 1), shape=(n, n))
    M = sp.csr_matrix((Mt[2], Mt[1] - 1, Mt[0] - 1), shape=(n, n))
    vol = 4.0 * 1.5 * 1.0
    assert np.max(np.abs(K @ np.ones(n))) < 1e-14
    assert abs(M.sum() - RHO_REF * CP * vol) < 1e-12 * RHO_REF * CP * vol
    x = mesh.xcoord[:, 0]
    assert abs(x @ (K @ x) - KAPPA * vol) < 1e-12
    assert abs(K - K.T).max() < 1e-15
    # distortion keeps conservation
    mesh2 = fem.Mesh(mesh.xcoord + rng.uniform(-0.05, 0.05, mesh.xcoord.shape), mesh.xconn)
    Kt2, _ = fem.assemble_t(mesh2, KAPPA, CP, RHO_REF)
    K2 = sp.csr_matrix((Kt2[2], Kt2[1] - 1, Kt2[0] - 1), shape=(n, n))
    assert np.max(np.abs(K2 @ np.ones(n))) < 1e-14


@pytest.mark.parametrize("tag,Tref", [("", 0.0), ("r", 4.0)])
def test_material_update_and_thermal_load_bit_for_bit(rod, golden, tag, Tref):
    from xtfem_amd import fem
    mesh = rod["mesh"]
    T = golden["th_T"][2 * mesh.nn:]
    ym, nu = fem.material_update(mesh, T, Tref, YM_REF, NU_REF)
    assert np.array_equal(ym, golden[f"th{tag}_ym"])
    assert np.array_equal(nu, golden[f"th{tag}_nu"])
    vv = fem.thermal_load(mesh, T, Tref, fem.ALPHA_REF, ym, nu)
    assert np.array_equal(vv, golden[f"th{tag}_vv"])
    K, M, *_ = fem.assemble(mesh, ym, nu, RHO_REF)
    assert digest(*K) == str(golden[f"th{tag}_K_digest"])
    assert digest(*M) == str(golden[f"th{tag}_M_digest"])


def test_material_update_law():
    """E0 - 1.6 dT - 0.1 dT^2 and nu0 + 5e-5 dT where T >= 0, unchanged below 0."""
    from xtfem_amd import fem
    mesh = fem.structured_hex(2, 1, 1, 2.0, 1.0, 1.0)
    for T0, expect in ((50.0, True), (-3.0, False)):
        ym, nu = fem.material_update(mesh, np.full(mesh.nn, T0), 0.0, 2.0e5, 0.3)
        if expect:
            assert np.allclose(ym, 2.0e5 - 1.6 * T0 - 0.1 * T0 * T0, rtol=1e-7)
            assert np.allclose(nu, 0.3 + 5e-5 * T0, rtol=1e-7)
        else:
            assert np.all(ym == 2.0e5) and np.all(nu == 0.3)


def test_thermal_load_free_expansion():
    """A uniform temperature rise dT on a body held only against rigid motion expands
    it freely: K u = vv gives u = alpha dT (x - x0), and vv is self-equilibrated."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from xtfem_amd import fem
    mesh = fem.structured_hex(3, 2, 2, 3.0, 2.0, 2.0)
    dT, alpha = 40.0, 1.2e-5
    ym, nu = fem.material_update(mesh, np.full(mesh.nn, dT), 0.0, 2.0e5, 0.3)
    vv = fem.thermal_load(mesh, np.full(mesh.nn, dT), 0.0, alpha, ym, nu)
    assert abs(vv[0::3].sum()) < 1e-9 * np.abs(vv).max()
    K, *_ = fem.assemble(mesh, ym, nu, 7.85e-9)
    n = 3 * mesh.nn
    A = sp.csr_matrix((K[2], K[1] - 1, K[0] - 1), shape=(n, n)).tolil()
    # 3-2-1 support at the origin corner, the x corner and the y corner
    x = mesh.xcoord
    o = int(np.flatnonzero((x == 0).all(axis=1))[0])
    ax = int(np.flatnonzero((x[:, 0] == 3.0) & (x[:, 1] == 0) & (x[:, 2] == 0))[0])
    ay = int(np.flatnonzero((x[:, 0] == 0) & (x[:, 1] == 2.0) & (x[:, 2] == 0))[0])
    fix = [3 * o, 3 * o + 1, 3 * o + 2, 3 * ax + 1, 3 * ax + 2, 3 * ay + 2]
    b = vv.copy()
    for d in fix:
        A[d, :] = 0.0
        A[:, d] = 0.0
        A[d, d] = 1.0
        b[d] = 0.0
    u = spl.spsolve(A.tocsr(), b)
    assert rel(u.reshape(-1, 3), alpha * dT * x) < 1e-9


def test_getfextc_t(golden):
    """Bit-level agreement with the reference's values, and the integral it closes:
    fextc(j) = int_{t1}^{t1+dt} dN_j/dt sin(w t) dt for the quadratic temporal shape
    functions, w = 2*3.14159/24 in FP32."""
    from scipy.integrate import quad
    from xtfem_amd import slab
    for (t1, dt), g in zip(golden["fxt_pts"], golden["fxt"]):
        f = slab.getfextc_t(t1, dt)
        assert rel(f, g) < 1e-15
        t2, tm = t1 + dt, t1 + dt / 2
        dN = [lambda t: 2 * ((t - t2) + (t - tm)) / dt ** 2,
              lambda t: -4 * ((t - t2) + (t - t1)) / dt ** 2,
              lambda t: 2 * ((t - tm) + (t - t1)) / dt ** 2]
        w = slab.W_THERMAL
        q = [quad(lambda t: dN[j](t) * np.sin(w * t), t1, t2, epsabs=1e-14)[0] for j in range(3)]
        assert rel(f, q) < 1e-8
    assert slab.W_THERMAL == float(np.float32(np.float32(6.28318) / np.float32(24)))


def test_heat_flux_and_fixed_temperature():
    """main3d.f90:429-478 on a plate: fnode from the incidences of one face, applied on
    both faces; fixed temperature on two planes."""
    from xtfem_amd import fem
    mesh = fem.structured_hex(4, 2, 3, 4.0, 2.0, 3.0, origin=(0.0, -1.0, 0.0))
    r = fem.heat_flux_load(mesh, 1, 1.0, [1.0, -1.0], 8.0)
    inc = (mesh.xcoord[mesh.xconn - 1][..., 1] == 1.0).sum()
    assert np.isclose(r.sum(), 2 * 8.0) and np.isclose(r.max(), 8.0 / inc * 4)
    assert np.all(r[np.abs(mesh.xcoord[:, 1]) < 1.0] == 0.0)
    fx = fem.fixed_temperature_mask(mesh, 2, [0.0, 3.0])
    assert fx.sum() == 2 * 5 * 3


def test_thermal_rhs_layout():
    from xtfem_amd import slab
    p = 5
    rext = np.arange(1.0, p + 1)
    fixed = np.zeros(3 * p, dtype=np.int32)
    fixed[[0, p, 2 * p]] = 1
    f = slab.thermal_rhs(p, [1.0, 2.0, 3.0], rext, jump=np.ones(3 * p),
                         init=np.full(3 * p, 0.5), bound=np.full(3 * p, 0.25), fixed=fixed)
    exp = np.concatenate([c * rext for c in (1.0, 2.0, 3.0)]) + 1.75
    exp[[0, p, 2 * p]] = 0.0
    assert np.array_equal(f, exp)


def test_rcm_renumbering_permutes_the_matrices_exactly():
    """Renumbering keeps element and local node order, so the assembled K of the
    renumbered mesh is P K P^T bit for bit; reverse Cuthill-McKee brings a scrambled
    numbering back to a narrow band."""
    import scipy.sparse as sp
    from xtfem_amd import fem
    mesh = fem.structured_hex(6, 4, 3, 6.0, 4.0, 3.0)
    rng = np.random.default_rng(11)
    scr = fem.renumber(mesh, rng.permutation(mesh.nn))
    assert fem.node_bandwidth(scr) > 3 * fem.node_bandwidth(mesh)
    perm = fem.rcm_order(scr)
    assert sorted(perm.tolist()) == list(range(mesh.nn))
    m2 = fem.renumber(scr, perm)
    # RCM's level sets are diagonal fronts: about twice the optimal lexicographic band
    assert fem.node_bandwidth(m2) <= 3 * fem.node_bandwidth(mesh)
    assert 2 * fem.node_bandwidth(m2) < fem.node_bandwidth(scr)
    K1, M1, *_ = fem.assemble(scr, 2.0e5, 0.3, 7.85e-9)
    K2, M2, *_ = fem.assemble(m2, 2.0e5, 0.3, 7.85e-9)
    n = 3 * mesh.nn
    P = np.repeat(3 * perm, 3) + np.tile(np.arange(3), mesh.nn)   # new DOF -> old DOF
    for A, B in ((K1, K2), (M1, M2)):
        a = sp.csr_matrix((A[2], A[1] - 1, A[0] - 1), shape=(n, n))
        b = sp.csr_matrix((B[2], B[1] - 1, B[0] - 1), shape=(n, n))
        ap = a[P][:, P].tocsr()
        ap.sort_indices()
        assert np.array_equal(ap.indptr, b.indptr) and np.array_equal(ap.indices, b.indices)
        assert np.array_equal(ap.data, b.data)
