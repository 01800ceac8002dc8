"""The autotuner's kernel variants (xt_tune_set knobs): every variant of the node-block
Kron-SpMM and of the banded-factor sweeps computes the same operator, checked against
the oracle's kronspmv (kronecker.f90:19-85) and the exact factor; the tuner leaves a
valid pick set and reuses it for the same problem."""
import numpy as np
import pytest

from conftest import rel, structured_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device, tune
    yield device
    for k in tune.KNOBS:                 # leave the built-in choices for later tests
        tune.set_knob(k, 0)


@pytest.mark.parametrize("R", [1, 5, 8, 32])
def test_spmm_variants_match_oracle(dev, oracle_lib, R):
    from xtfem_amd import slab, tune
    S = structured_system(12, 5, 4, 24.0, 10.0, 8.0)
    op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
    assert op.format == "blocks"
    rng = np.random.default_rng(R)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * rng.uniform(10, 100, R))
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    X = rng.standard_normal((R, 6 * S["p"]))
    ref = np.stack([oracle_lib.kronspmv(A[r], B[r], S["Kb"], S["Mb"], X[r]) for r in range(R)])
    splits = (4, 8) if R == 1 else (0,)
    for unr in (2, 4, 8):
        for wpb in (1, 2, 4):
            for spl in splits:
                for lds in ((0,) if R == 1 else (1, 2)):
                    tune.set_knob("spmm_unr", unr)
                    tune.set_knob("spmm_wpb", wpb)
                    tune.set_knob("spmm_r1split", spl)
                    tune.set_knob("spmm_lds", lds)
                    for fp32, tol in ((False, 1e-12), (True, 2e-5)):
                        Y = op.apply_host(A, B, X, fp32=fp32)
                        for r in range(R):
                            assert rel(Y[r], ref[r]) < tol, (unr, wpb, spl, lds, fp32, r)
    for k in tune.KNOBS:
        tune.set_knob(k, 0)


def test_band_tile_variants_and_autotune(dev, oracle_lib):
    """The 64- and 128-row tiles of the FP32 batched passes against the exact factor, then
    a full autotune: picks among the candidates, knobs left set, cached on reuse."""
    import torch
    from xtfem_amd import slab, tune
    S = structured_system(130, 1, 1, 130.0, 1.0, 1.0)
    pr = dev.BandPreconditioner(S["Kb"], block=128)
    op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
    p, R = S["p"], 32
    alu, jlu, ju, ierr = oracle_lib.ilut(S["Kb"], p, 0.0)
    rng = np.random.default_rng(3)
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * rng.uniform(10, 100, R))
    A = c["k"] + c["ktn"]
    X = rng.standard_normal((R, 6 * p))
    ref = np.stack([oracle_lib.kronlusol(A[r], alu, jlu, ju, X[r]) for r in range(R)])
    minv = torch.as_tensor(np.linalg.inv(A), device="cuda").float().contiguous()
    x = dev.to_interleaved(X, p, 6).float().contiguous()
    for tile in (64, 128):
        tune.set_knob("band_tile", tile)
        y = torch.empty_like(x)
        pr.apply(minv, x, y, R, fp32=True)
        Y = dev.from_interleaved(y.double().contiguous(), p, 6, R)
        for r in range(R):
            assert rel(Y[r], ref[r]) < 1e-3, (tile, r)
    out = tune.autotune(op, pr, R, fp32=True, reps=3)
    assert out["picks"]["spmm_unr"] in (2, 4, 8) and out["picks"]["band_tile"] in (64, 128)
    for k, v in out["picks"].items():
        assert tune.get_knob(k) == v
    again = tune.autotune(op, pr, R, fp32=True)
    assert again.get("cached") and again["picks"] == out["picks"]


@pytest.mark.parametrize("block", [64, 256])
def test_band_chain_and_level_variants(dev, oracle_lib, block):
    """The chain steps as split-k tile products (XT_TUNE_BAND_CHAIN 2) or one workgroup
    per 16 x 16 tile (1), with factors built for 0, 1, 2 and up to 4 pairing levels
    (XT_TUNE_BAND_LEVELS), in FP64 and FP32 at 6, 30 and 192 columns: every variant
    applies the exact factor (the oracle's ilut at full fill, ilut.f:2169-2170)."""
    import torch
    from xtfem_amd import fem, slab, tune
    # a 260 x 1 x 1 strip clamped along its long side y = 0 (well conditioned: the FP64
    # factors agree to rounding), 49 blocks of 64 rows or 13 of 256
    mesh = fem.structured_hex(260, 1, 1, 260.0, 1.0, 1.0)
    K, M, *_ = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    S = dict(Kb=fem.applybc(K, fem.clamp_mask(mesh, 1, [0.0])), p=3 * mesh.nn)
    p = S["p"]
    alu, jlu, ju, ierr = oracle_lib.ilut(S["Kb"], p, 0.0)
    assert ierr == 0
    rng = np.random.default_rng(block)
    try:
        for levels in (1, 2, 3, 5):
            tune.set_knob("band_levels", levels)
            pr = dev.BandPreconditioner(S["Kb"], block=block)
            assert pr.block == block
            for R in (1, 5, 32):
                c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * rng.uniform(10, 100, R))
                A = c["k"] + c["ktn"]
                X = rng.standard_normal((R, 6 * p))
                ref = np.stack([oracle_lib.kronlusol(A[r], alu, jlu, ju, X[r])
                                for r in range(R)])
                for chain in (1, 2):
                    tune.set_knob("band_chain", chain)
                    for fp32, tol in ((False, 1e-9), (True, 1e-3)):
                        dt = torch.float32 if fp32 else torch.float64
                        minv = torch.as_tensor(np.linalg.inv(A), device="cuda").to(dt).contiguous()
                        x = dev.to_interleaved(X, p, 6).to(dt).contiguous()
                        y = torch.empty_like(x)
                        pr.apply(minv, x, y, R, fp32=fp32)
                        Y = dev.from_interleaved(y.double().contiguous(), p, 6, R)
                        for r in range(R):
                            assert rel(Y[r], ref[r]) < tol, (levels, R, chain, fp32, r)
            del pr
    finally:
        tune.set_knob("band_levels", 0)
        tune.set_knob("band_chain", 0)


@pytest.mark.parametrize("smallc", [1, 2, 3])
def test_band_few_column_variants(dev, oracle_lib, smallc):
    """Sweeps with 6 and 12 columns on 256-row blocks (13 blocks: batched passes of several
    blocks, one pairing level): the VALU rows kernel (1), the matrix-core kernels with
    64-column tiles (2) and with 16-column tiles in the batched passes (3), FP64 and
    FP32, with and without the pairing level, against the exact factor
    (ilut.f:2169-2170)."""
    import torch
    from xtfem_amd import fem, slab, tune
    mesh = fem.structured_hex(260, 1, 1, 260.0, 1.0, 1.0)
    K, M, *_ = fem.assemble(mesh, 2.0e5, 0.3, 7.85e-9)
    Kb, p = fem.applybc(K, fem.clamp_mask(mesh, 1, [0.0])), 3 * mesh.nn
    alu, jlu, ju, ierr = oracle_lib.ilut(Kb, p, 0.0)
    assert ierr == 0
    rng = np.random.default_rng(smallc)
    pr = dev.BandPreconditioner(Kb, block=256)
    assert pr.block == 256 and pr.nblocks == 13
    try:
        tune.set_knob("band_smallc", smallc)
        for pair in (1, 2):
            tune.set_knob("band_pair", pair)
            for R in (1, 2):
                c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * rng.uniform(10, 100, R))
                A = c["k"] + c["ktn"]
                X = rng.standard_normal((R, 6 * p))
                ref = np.stack([oracle_lib.kronlusol(A[r], alu, jlu, ju, X[r]) for r in range(R)])
                for fp32, tol in ((False, 1e-9), (True, 1e-3)):
                    dt = torch.float32 if fp32 else torch.float64
                    minv = torch.as_tensor(np.linalg.inv(A), device="cuda").to(dt).contiguous()
                    x = dev.to_interleaved(X, p, 6).to(dt).contiguous()
                    y = torch.empty_like(x)
                    pr.apply(minv, x, y, R, fp32=fp32)
                    Y = dev.from_interleaved(y.double().contiguous(), p, 6, R)
                    for r in range(R):
                        assert rel(Y[r], ref[r]) < tol, (smallc, pair, R, fp32, r)
    finally:
        tune.set_knob("band_smallc", 0)
        tune.set_knob("band_pair", 0)
