"""GPU parity of the exact banded-factor preconditioner (block sweeps) and of slab
solves that use it.

The oracle for K^-1 is ilut with lfil >= n and droptol = 0, which is the exact LU
factor (ilut.f:2169-2170; the oracle's ilut is pinned bit for bit against the
compiled reference), applied by the oracle's kronlusol.  Tolerances: FP64 sweeps
agree with it to 1e-9 relative (explicitly inverted diagonal blocks); FP32 sweeps to
1e-3 relative (FP32 factor of a matrix with condition number ~1e5); slab solves to the
true relative residual 1e-9 and 1e-6 relative displacement against the oracle's FP64
GMRES with the exact factor converged to 1e-12.
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


def notched_plate(nx, ny, nz, notch_len, order="zyx"):
    """Plate clamped at y = 0 with an edge notch of notch_len elements at mid-height."""
    from xtfem_amd import fem
    keep = np.ones((nx, ny, nz), bool)
    keep[0:notch_len, ny // 2, :] = False
    mesh = fem.structured_hex(nx, ny, nz, float(nx), float(ny), float(nz), keep=keep,
                              order=order)
    K, M, *_ = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    bc = fem.clamp_mask(mesh, 1, [0.0])
    return dict(mesh=mesh, Kb=fem.applybc(K, bc), Mb=fem.applybc(M, bc), bc=bc,
                p=3 * mesh.nn)


def exact_ilut(oracle_lib, K):
    p = len(K[0]) - 1
    alu, jlu, ju, ierr = oracle_lib.ilut(K, p, 0.0)
    assert ierr == 0
    return alu, jlu, ju


@pytest.mark.parametrize("shape,block,R", [((20, 2, 2, 0), 0, 3),     # many 64-row blocks, c = 18
                                           ((20, 2, 2, 0), 0, 2),     # c = 12: 16-column rows kernel
                                           ((12, 9, 2, 4), 0, 12),    # c = 72: column tiles
                                           ((12, 9, 2, 4), 0, 32),    # c = 192
                                           ((12, 9, 2, 4), 256, 1),   # one padded block, c = 6
                                           ((260, 1, 1, 0), 64, 32),  # 49 blocks: 3 pairing levels
                                           ((260, 1, 1, 0), 64, 2),   # the same, rows kernel
                                           ((520, 1, 1, 0), 128, 32)])  # FP32 128-row tiles
def test_band_sweeps_match_exact_lu(dev, oracle_lib, shape, block, R):
    from xtfem_amd import slab
    S = notched_plate(*shape)
    p, K = S["p"], S["Kb"]
    pr = dev.BandPreconditioner(K, block=block)
    assert pr.block % 64 == 0 and pr.block >= pr.bandwidth
    assert pr.nblocks == -(-p // pr.block)
    if shape[0] == 20:
        assert pr.nblocks > 4 and p % pr.block != 0
    if shape[0] == 260:   # recurrences of 48 steps: paired over three levels (chain of 6)
        assert pr.nblocks >= 49
    if shape[0] == 520:   # 128-row blocks, >= 23 blocks in a batched pass: the FP32
        assert pr.block == 128 and pr.nblocks >= 49   # sweeps take the 128 x 64 tiles
    alu, jlu, ju = exact_ilut(oracle_lib, K)
    rng = np.random.default_rng(7 + R)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * rng.uniform(10, 100, R))
    A = c["k"] + c["ktn"]
    X = rng.standard_normal((R, 6 * p))
    import torch
    minv = torch.as_tensor(np.linalg.inv(A), device="cuda").contiguous()
    x = dev.to_interleaved(X, p, 6)
    ref = np.stack([oracle_lib.kronlusol(A[r], alu, jlu, ju, X[r]) for r in range(R)])
    for fp32, tol in ((False, 1e-9), (True, 1e-3)):
        dt = torch.float32 if fp32 else torch.float64
        xi = x.to(dt).contiguous()
        y = torch.empty_like(xi)
        pr.apply(minv.to(dt).contiguous(), xi, y, R, fp32=fp32)
        Y = dev.from_interleaved(y.to(torch.float64).contiguous(), p, 6, R)
        for r in range(R):
            assert rel(Y[r], ref[r]) < tol, (fp32, r, rel(Y[r], ref[r]))
    # without the temporal mix: K y_j = x_j column by column
    y = torch.empty_like(x)
    pr.apply(None, x, y, R)
    Y = dev.from_interleaved(y, p, 6, R)
    for r in range(R):
        Kx = np.concatenate([oracle_lib.amux(*K, Y[r, j * p:(j + 1) * p]) for j in range(6)])
        assert rel(Kx, X[r]) < 1e-9


def test_band_rejects_indefinite(dev):
    """An unconstrained K (rigid modes) is singular: XT_E_SINGULAR, no crash."""
    from xtfem_amd import fem
    from xtfem_amd.api import XtError
    mesh = fem.structured_hex(3, 2, 2, 3.0, 2.0, 2.0)
    K = fem.assemble(mesh, 2.0e5, 0.3, 1.0)[0]
    K = (K[0], K[1], -K[2])                          # negative definite
    with pytest.raises(XtError, match="code -3"):
        dev.BandPreconditioner(K)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_slab_solve_with_band_factor(dev, oracle_lib, mode):
    from xtfem_amd import fem, slab
    S = notched_plate(16, 10, 2, 5)
    p = S["p"]
    R = 4
    rng = np.random.default_rng(20261202)
    amps, freqs = rng.uniform(20, 120, R), rng.uniform(10, 100, R)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * freqs)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    ytop = S["mesh"].xcoord[:, 1].max()
    rext0 = fem.traction_load(S["mesh"], 1, ytop, 1, 1.0)
    fx, _ = slab.getfextc_m(0.0, 2 * np.pi * freqs, 0.5)
    rhs = np.array([slab.slab_rhs(p, fx[r], None, amps[r] * rext0, fixed=np.tile(S["bc"], 6))
                    for r in range(R)])
    op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
    pr = dev.BandPreconditioner(S["Kb"])
    x, its, ierr, relres = dev.slab_solve(op, pr, A, B, rhs, im=30, eps=1e-10, maxits=200,
                                          mode=mode)
    assert np.all(ierr == 0), (ierr, its)
    assert np.all(its <= 12), its
    prec = exact_ilut(oracle_lib, S["Kb"])
    for r in range(R):
        y = oracle_lib.kronspmv(A[r], B[r], S["Kb"], S["Mb"], x[r])
        rr = np.linalg.norm(rhs[r] - y) / np.linalg.norm(rhs[r])
        assert rr < 1e-9 and relres[r] < 1e-9, (r, rr, relres[r])
        ref, ie, _, _ = oracle_lib.kronpgmres(A[r], B[r], S["Kb"], S["Mb"], prec, rhs[r], im=30,
                                              eps=1e-12, maxits=100)
        assert ie == 0
        assert rel(x[r], ref) < 1e-6


@pytest.mark.parametrize("R,fp32", [(1, False), (1, True), (8, True), (32, False)])
def test_band_sweeps_graph_replay(dev, R, fp32):
    """The sweeps replayed from a captured HIP graph equal the launch-by-launch sweeps
    bit for bit, on every replay with eager work and synchronisations between them,
    and after a refactorization (new factor buffers: a new capture)."""
    import torch
    from xtfem_amd._lib import lib
    S = notched_plate(12, 9, 2, 4)
    p, K = S["p"], S["Kb"]
    pr = dev.BandPreconditioner(K)
    assert pr.nblocks > 2
    dt = torch.float32 if fp32 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(5 + R)
    x = (torch.rand(p * 6 * R, device="cuda", generator=g, dtype=torch.float64) - 0.5).to(dt)

    def apply(on):
        lib().xt_graphs_enable(int(on))
        y = torch.empty_like(x)
        pr.apply(None, x, y, R, q=6, fp32=fp32)
        torch.cuda.synchronize()
        return y

    try:
        want = apply(False)
        for rep in range(4):
            got = apply(True)
            assert torch.equal(got, want), rep
            z = torch.sin(x) * 3.0        # eager work between replays
            torch.cuda.synchronize()
            assert z.numel() == x.numel()
        # refactor from the device CSR with the same values: new buffers, same result
        rp = torch.as_tensor(np.asarray(K[0], dtype=np.int32), device="cuda")
        ci = torch.as_tensor(np.asarray(K[1], dtype=np.int32), device="cuda")
        va = torch.as_tensor(np.asarray(K[2], dtype=np.float64), device="cuda")
        pr.refactor(rp, ci, va)
        want2 = apply(False)
        for rep in range(3):
            assert torch.equal(apply(True), want2), rep
        assert rel(want2.double().cpu().numpy(), want.double().cpu().numpy()) < 1e-12
    finally:
        lib().xt_graphs_enable(1)


def test_band_factor_after_renumbering_and_wide_blocks(dev):
    """A scrambled mesh renumbered by reverse Cuthill-McKee takes the band factor; and a
    cross-section of 31 x 31 nodes (half-bandwidth 2885 DOFs, blocks of 2944 rows, above
    the former 2048 limit) factors and solves K y = x to FP64 accuracy."""
    import scipy.sparse as sp
    import torch
    from xtfem_amd import fem
    base = fem.structured_hex(10, 6, 3, 10.0, 6.0, 3.0)
    scr = fem.renumber(base, np.random.default_rng(3).permutation(base.nn))
    cases = [fem.renumber(scr, fem.rcm_order(scr)),
             fem.structured_hex(6, 30, 30, 6.0, 30.0, 30.0)]
    for mesh in cases:
        K, *_ = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
        bc = fem.clamp_mask(mesh, 0, [0.0])
        Kb = fem.applybc(K, bc)
        p = 3 * mesh.nn
        pr = dev.BandPreconditioner(Kb)
        assert pr.block >= pr.bandwidth
        rng = np.random.default_rng(p)
        X = rng.standard_normal((6, p))
        x = torch.as_tensor(np.ascontiguousarray(X.T).ravel(), device="cuda")
        y = torch.empty_like(x)
        pr.apply(None, x, y, 1, q=6, fp32=False)
        Y = y.cpu().numpy().reshape(p, 6).T
        A = sp.csr_matrix((Kb[2], Kb[1] - 1, Kb[0] - 1), shape=(p, p))
        for j in range(6):
            assert np.linalg.norm(A @ Y[j] - X[j]) / np.linalg.norm(X[j]) < 1e-8, (p, j)


@pytest.mark.parametrize("where", ["block1", "middle", "boundary", "last", "full"])
def test_band_refresh_matches_fresh_factor(dev, oracle_lib, where):
    """xt_prec_refresh_band: K changed only in rows and columns >= r0 (K' = K + 0.3 P K22
    P^T, the trailing principal part scaled: still SPD); the factor of K refreshed from
    r0 on gives the sweeps of a factor of K' built afresh (FP64 to 1e-10, FP32 to 1e-4,
    with and without the pairing levels) and solves K' y = x to 1e-9, on a 49-block
    factor with three pairing levels.  Rows r0 in block 1 (E_0 from the kept D_0), in a
    middle block, at a block boundary, in the last block, and r0 = 0 (whole refactor)."""
    import torch
    S = notched_plate(260, 1, 1, 0)
    p, K = S["p"], S["Kb"]
    ia, ja, a = (np.asarray(v) for v in K)
    b = 64
    r0 = {"block1": b + 10, "middle": 23 * b + 37, "boundary": 30 * b,
          "last": p - 5, "full": 0}[where]
    rows = np.repeat(np.arange(p), np.diff(ia))
    a2 = a * np.where((rows >= r0) & (ja - 1 >= r0), 1.3, 1.0)
    K2 = (ia, ja, a2)
    pr = dev.BandPreconditioner(K, block=b)
    assert pr.block == b and pr.nblocks >= 49
    dv = lambda v, t: torch.as_tensor(np.ascontiguousarray(v, dtype=t), device="cuda")
    pr.refactor(dv(ia, np.int32), dv(ja, np.int32), dv(a2, np.float64), first_row=r0)
    fresh = dev.BandPreconditioner(K2, block=b)
    rng = np.random.default_rng(11)
    for R in (2, 32):                      # FP64: paired (<= 64 columns) and unpaired
        X = rng.standard_normal((R, 6 * p))
        x = dev.to_interleaved(X, p, 6)
        for fp32, tol in ((False, 1e-10), (True, 1e-4)):
            dt = torch.float32 if fp32 else torch.float64
            xi = x.to(dt).contiguous()
            y1, y2 = torch.empty_like(xi), torch.empty_like(xi)
            pr.apply(None, xi, y1, R, fp32=fp32)
            fresh.apply(None, xi, y2, R, fp32=fp32)
            Y1, Y2 = (dev.from_interleaved(y.double().contiguous(), p, 6, R) for y in (y1, y2))
            for r in range(R):
                assert rel(Y1[r], Y2[r]) < tol, (where, R, fp32, r, rel(Y1[r], Y2[r]))
            if not fp32:
                for r in range(0, R, 7):
                    Kx = np.concatenate([oracle_lib.amux(ia, ja, a2, Y1[r, j * p:(j + 1) * p])
                                         for j in range(6)])
                    assert rel(Kx, X[r]) < 1e-9, (where, R, r)
