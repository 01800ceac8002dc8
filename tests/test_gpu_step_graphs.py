"""Arnoldi steps replayed from captured HIP graphs (the slab solver's step graphs): every
solve equals the launch-by-launch solve bit for bit, over repeated solves (captures on
the second use of a step, replays after) with eager work and synchronisations between
them, for the ILUT path (the rod, FP64, one load case) and the banded factor with the
split preconditioner (a plate, FP32-IR, several load cases), for the segmented
multi-system solve, and for one-case solves whose splits map the case to different
shifted factors."""
import numpy as np
import pytest

from conftest import structured_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


def _rhs(S, freq, amp):
    from xtfem_amd import fem, slab
    w = 2 * np.pi * np.asarray(freq)
    dt = np.round(0.5 * np.asarray(freq)) / np.asarray(freq)
    c = slab.stcoeff_m(0.0, dt, w)
    A, B = c["k"] + c["ktn"], c["m"] + c["mtn"]
    x = S["mesh"].xcoord
    rext = fem.traction_load(S["mesh"], 0, float(x[:, 0].max()), 0, 1.0)
    fx, _ = slab.getfextc_m(0.0, w, dt)
    rhs = np.array([slab.slab_rhs(S["p"], fx[r] * amp[r], None, rext, fixed=np.tile(S["bc"], 6))
                    for r in range(len(freq))])
    return A, B, rhs


@pytest.mark.parametrize("case", ["ilut_fp64", "band_split_fp32ir", "segmented"])
def test_step_graphs_bit_identical(dev, case):
    import torch
    from xtfem_amd._lib import lib
    if case == "ilut_fp64":
        S = structured_system(4, 4, 12, 10.0, 10.0, 30.0)
        A, B, rhs = _rhs(S, [50.0], [100.0])
        op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
        pr = dev.IlutPreconditioner(S["Kb"], 20, 1e-4)
        mode, xtol = 0, 0.0
    else:
        S = structured_system(16, 4, 3, 32.0, 8.0, 6.0, rho=50 * 7.85e-9)
        A, B, rhs = _rhs(S, [14.0, 47.0, 88.0], [50.0, 90.0, 30.0])
        op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
        pr0 = dev.BandPreconditioner(S["Kb"])
        mu = dev.lowest_mode(pr0, S["Mb"], S["bc"])
        pr = dev.BandPreconditioner(S["Kb"], M=S["Mb"],
                                    shifts=dev.shifts_for_cases(dev.dominant_temporal_mode(A, B), mu))
        mode, xtol = 1, 1e-7
    R, p = len(rhs), S["p"]
    b = dev.to_interleaved(rhs, p, 6)

    def solve(on):
        lib().xt_graphs_enable(int(on))
        x = torch.zeros_like(b)
        if case == "segmented":
            seg = np.array([0, p // 6 * 3, p])            # two segments (the path, not the physics)
            out = dev.slab_solve_seg(op, pr, seg, A, B, b, x, R, mode=mode, xtol=xtol)
        else:
            out = dev.slab_solve_device(op, pr, A, B, b, x, R, mode=mode, xtol=xtol)
        torch.cuda.synchronize()
        return x.clone(), np.asarray(out[1]).ravel()

    from xtfem_amd import tune
    tune.set_knob("ir_skip32", 1)        # every solve probes FP32 first: all solves alike
    try:
        want, its0 = solve(False)
        for rep in range(4):
            got, its = solve(True)
            assert torch.equal(got, want), (case, rep)
            assert np.array_equal(its, its0)
            z = torch.cos(b) * 2.0                        # eager work between solves
            torch.cuda.synchronize()
            assert z.numel() == b.numel()
    finally:
        lib().xt_graphs_enable(1)
        tune.set_knob("ir_skip32", 0)


def test_step_graphs_across_precisions_and_reallocation(dev):
    """Solves that alternate FP32-IR, FP64 GMRES and FP64 FGMRES with FP32 sweeps on one
    operator and banded factor (the FP32 and FP64 sweeps share the factor's workspace),
    with another factor built and destroyed between them (device memory allocated and
    freed while step graphs are cached): every solve equals its launch-by-launch
    counterpart bit for bit."""
    import torch
    from xtfem_amd._lib import lib
    S = structured_system(16, 4, 3, 32.0, 8.0, 6.0)
    A, B, rhs = _rhs(S, [14.0, 47.0], [50.0, 90.0])
    op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
    pr = dev.BandPreconditioner(S["Kb"])
    R, p = len(rhs), S["p"]
    b = dev.to_interleaved(rhs, p, 6)

    def solve(on, mode):
        lib().xt_graphs_enable(int(on))
        x = torch.zeros_like(b)
        out = dev.slab_solve_device(op, pr, A, B, b, x, R, mode=mode,
                                    xtol=1e-7 if mode == 1 else 0.0)
        torch.cuda.synchronize()
        return x.clone(), np.asarray(out[1]).ravel()

    from xtfem_amd import tune
    tune.set_knob("ir_skip32", 1)        # every solve probes FP32 first: all solves alike
    try:
        want = {m: solve(False, m) for m in (1, 0, 2)}
        for rep in range(3):
            for m in (1, 0, 2, 1):
                got, its = solve(True, m)
                assert torch.equal(got, want[m][0]), (rep, m)
                assert np.array_equal(its, want[m][1]), (rep, m)
            other = dev.BandPreconditioner(S["Kb"], block=128)   # allocate, then free
            del other
            torch.cuda.synchronize()
    finally:
        lib().xt_graphs_enable(1)
        tune.set_knob("ir_skip32", 0)


def test_step_graphs_follow_each_solves_split(dev):
    """One-case FP64 solves one after another on one operator and split preconditioner,
    the cases' dominant temporal modes going to different shifted factors: the Arnoldi
    step graphs are keyed on the split's groups, so a graph captured for one case's
    split is never replayed for another's.  Each solve reaches the 1e-9 residual and
    equals its launch-by-launch counterpart bit for bit, over repeated rounds."""
    import torch
    from xtfem_amd._lib import lib
    S = structured_system(16, 4, 3, 32.0, 8.0, 6.0, rho=50 * 7.85e-9)
    A, B, rhs = _rhs(S, [14.0, 47.0, 88.0], [50.0, 90.0, 30.0])
    op = dev.KronOperator([S["Kb"], S["Mb"]], q=6)
    pr0 = dev.BandPreconditioner(S["Kb"])
    mu = dev.lowest_mode(pr0, S["Mb"], S["bc"])
    pr = dev.BandPreconditioner(S["Kb"], M=S["Mb"],
                                shifts=dev.shifts_for_cases(dev.dominant_temporal_mode(A, B), mu))
    assert len(pr.shifts) >= 2, pr.shifts
    p = S["p"]

    def solve(on, r):
        lib().xt_graphs_enable(int(on))
        b = dev.to_interleaved(rhs[r:r + 1], p, 6)
        x = torch.zeros_like(b)
        out = dev.slab_solve_device(op, pr, A[r:r + 1], B[r:r + 1], b, x, 1, eps=1e-9, mode=0)
        torch.cuda.synchronize()
        return x.clone(), float(np.asarray(out[3]).ravel()[0])

    try:
        want = [solve(False, r) for r in range(3)]
        for rep in range(3):
            for r in range(3):
                x, rr = solve(True, r)
                assert rr <= 1e-9, (rep, r, rr)
                assert torch.equal(x, want[r][0]), (rep, r)
    finally:
        lib().xt_graphs_enable(1)
