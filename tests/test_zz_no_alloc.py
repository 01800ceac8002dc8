"""The allocation-free ILUT preconditioner apply (xt_prec_apply_noalloc) against
xt_prec_apply, through the C ABI: the same kernels on the same inputs, so the results
are bit-identical.  Two calls in a row, the second with fewer right-hand sides, in FP64
and FP32, with and without the temporal mix; then the error paths (size above the
fixed scratch, banded factor)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from xtfem_amd import device
    return device


def test_noalloc_matches_prec_apply(dev, rod):
    import torch
    from xtfem_amd import slab
    from xtfem_amd._lib import check, lib
    p = rod["p"]
    pr = dev.IlutPreconditioner(rod["Kb"], 20, 1e-4)
    rng = np.random.default_rng(3)
    s = torch.cuda.current_stream().cuda_stream
    for R in (5, 2):                                   # the second call is smaller
        freqs = rng.uniform(10, 100, R)
        c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * freqs)
        minv64 = torch.as_tensor(np.linalg.inv(c["k"] + c["ktn"]), device="cuda").contiguous()
        x64 = torch.as_tensor(rng.standard_normal(p * 6 * R), device="cuda")
        for fp32 in (False, True):
            dt = torch.float32 if fp32 else torch.float64
            x = x64.to(dt).contiguous()
            for minv in (minv64.to(dt).contiguous(), None):
                mp = minv.data_ptr() if minv is not None else None
                y0 = torch.full_like(x, float("nan"))
                y1 = torch.full_like(x, float("nan"))
                check(lib().xt_prec_apply(pr.h, R, 6, mp, x.data_ptr(), y0.data_ptr(), int(fp32),
                                          s), "prec_apply")
                check(lib().xt_prec_apply_noalloc(pr.h, R, 6, mp, x.data_ptr(), y1.data_ptr(),
                                                  int(fp32), s), "prec_apply_noalloc")
                torch.cuda.synchronize()
                assert torch.isfinite(y0).all()
                assert torch.equal(y0, y1), (R, fp32, minv is None)


def test_noalloc_rejects_what_it_cannot_stage(dev, rod):
    import torch
    from xtfem_amd._lib import lib
    p = rod["p"]
    pr = dev.IlutPreconditioner(rod["Kb"], 20, 1e-4)
    cap = 16 << 20                                      # XT_NOALLOC_SCRATCH_BYTES
    R = cap // (p * 6 * 8) + 1                          # one right-hand side too many
    x = torch.zeros(p * 6 * R, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    minv = torch.zeros((R, 6, 6), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    rc = lib().xt_prec_apply_noalloc(pr.h, R, 6, minv.data_ptr(), x.data_ptr(), y.data_ptr(), 0, s)
    assert rc == -1 and b"scratch" in lib().xt_last_error()
    band = dev.BandPreconditioner(rod["Kb"])
    rc = lib().xt_prec_apply_noalloc(band.h, 1, 6, None, x.data_ptr(), y.data_ptr(), 0, s)
    assert rc == -5
