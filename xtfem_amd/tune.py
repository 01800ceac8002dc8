"""Autotuner of the kernel variants (xt_tune_set knobs) on the caller's own operator and
preconditioner.

For a given operator, preconditioner and load-case count R it times, with HIP events on
the launching stream (the library's per-launch profiler), the candidates of

* the node-block Kron-SpMM: entries in flight per pass (2, 4, 8), waves per workgroup
  (1, 2, 4), for one load case lanes per node row (4, 8), for several the kernel that
  stages each slice's neighbour rows of X in LDS (off, on);
* the banded-factor sweeps: the tile rows of the batched passes (64, 128 in FP32), and
  for a few columns the VALU rows kernel or the matrix-core kernels (64- or 16-column
  tiles in the batched passes);
* the slab solve, timed on whole solves when a solve is given: the least residual
  reduction an FP32 refinement step of GMRES-IR must reach for a case to stay in FP32
  (2, 8, 32, 128), the iterations an FP32 inner cycle gets before a case moves to FP64
  cycles (1, 2, 4, 8) and the Arnoldi steps between host convergence checks (1, 2, 4);

keeps the fastest of each (knobs tuned one after another, the earlier ones fixed at their
pick) and leaves them set.  The choice is cached per problem signature (p, nonzeros,
band blocks, R, precision), in memory and, if XTFEM_TUNE_CACHE names a file, in that JSON
file, so a process that solves the same problem again skips the timing.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np
import torch

from ._lib import check, lib

KNOBS = {"spmm_unr": 0, "spmm_wpb": 1, "spmm_r1split": 2, "band_tile": 3, "gmres_check": 4,
         "spmm_lds": 5, "band_smallc": 6, "ir_fp32_cycle": 7, "ir_min_gain": 8,
         "ir_skip32": 9, "band_chain": 10, "band_levels": 11,
         "gmres_reorth": 12, "gmres_compact": 13, "band_chain_few": 14,
         "band_pair": 15}
XT_PROF_SPMM, XT_PROF_BAND, XT_PROF_BAND64 = 0, 1, 4
_CACHE: dict = {}


def set_knob(name, value):
    check(lib().xt_tune_set(KNOBS[name], int(value)), f"tune_set {name}")


def get_knob(name):
    return int(lib().xt_tune_get(KNOBS[name]))


def _timed(kind, fn, reps):
    """Median per-call kernel time (ms) of fn() over reps calls, HIP events of `kind`."""
    times = []
    for _ in range(reps):
        check(lib().xt_prof_enable(1 << kind), "prof_enable")
        fn()
        ms, n = C.c_double(), C.c_longlong()
        check(lib().xt_prof_read(kind, C.addressof(ms), C.addressof(n), None, None), "prof_read")
        times.append(ms.value)
    check(lib().xt_prof_enable(0), "prof_enable")
    return float(np.median(times))


def _pick(name, cands, kind, fn, reps):
    res = {}
    for v in cands:
        set_knob(name, v)
        fn()                                  # warm-up (first use of a variant)
        res[v] = _timed(kind, fn, reps)
    best = min(res, key=res.get)
    set_knob(name, best)
    return best, res


def _wall(fn, reps):
    import time
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        times.append(1e3 * (time.perf_counter() - t0))
    return float(np.median(times))


def signature(op, pr, R, fp32):
    band = (pr.block, pr.nblocks) if getattr(pr, "block", None) else (0, 0)
    return f"p{op.p}-nnz{sum(op.nnz)}-b{band[0]}x{band[1]}-R{R}-{'f32' if fp32 else 'f64'}"


def _load_file():
    path = os.environ.get("XTFEM_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _save_file(entry_key, entry):
    path = os.environ.get("XTFEM_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data[entry_key] = entry
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def autotune(op, pr, R, fp32=True, reps=7, solve=None):
    """Tune the knobs for this operator / preconditioner / R and leave them set.

    solve: optional zero-argument callable running one slab solve with these handles
    (then the host-check interval is tuned on whole solves).  Returns a dict with the
    picks and the timings of every candidate (ms per call)."""
    key = signature(op, pr, R, fp32)
    cached = _CACHE.get(key) or _load_file().get(key)
    if cached:
        for k, v in cached["picks"].items():
            set_knob(k, v)
        return dict(cached, cached=True)
    for k in KNOBS:
        set_knob(k, 0)
    dt = torch.float32 if fp32 else torch.float64
    p, q = op.p, op.q
    g = torch.Generator(device="cuda").manual_seed(1)
    x = (torch.rand(p * q * R, device="cuda", generator=g, dtype=dt) * 2 - 1).contiguous()
    y = torch.empty_like(x)
    from . import slab
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * 20.0)
    T = op.tmat(np.tile(np.stack([c["k"] + c["ktn"], c["m"] + c["mtn"]])[None], (R, 1, 1, 1)), fp32)
    spmm = lambda: op.apply(T, x, y, R, fp32=fp32)
    out = {"signature": key, "picks": {}, "timings_ms": {}}
    if op.format == "blocks":
        for name, cands in (("spmm_unr", (2, 4, 8)), ("spmm_wpb", (1, 2, 4))) + \
                ((("spmm_r1split", (4, 8)),) if R == 1 else (("spmm_lds", (1, 2)),)):
            best, res = _pick(name, cands, XT_PROF_SPMM, spmm, reps)
            out["picks"][name] = best
            out["timings_ms"][name] = res
    if getattr(pr, "block", None):
        minv = torch.as_tensor(np.linalg.inv(c["k"] + c["ktn"]), device="cuda").to(dt)
        minv = minv.expand(R, q, q).contiguous()
        sweep = lambda: pr.apply(minv, x, y, R, fp32=fp32)
        cands = (64, 128) if fp32 and pr.block % 128 == 0 else (64,)
        kb = XT_PROF_BAND if fp32 else XT_PROF_BAND64
        best, res = _pick("band_tile", cands, kb, sweep, max(3, reps // 2))
        out["picks"]["band_tile"] = best
        out["timings_ms"]["band_tile"] = res
        # few columns (a shifted factor's group, one load case): rows kernel or matrix cores
        xs = x[:p * q].contiguous()                 # one load case: q columns
        ys = torch.empty_like(xs)
        sweep_small = lambda: pr.apply(None, xs, ys, 1, q=q, fp32=fp32)
        best, res = _pick("band_smallc", (1, 2, 3), kb, sweep_small, max(3, reps // 2))
        out["picks"]["band_smallc"] = best
        out["timings_ms"]["band_smallc"] = res
    if solve is not None:
        for name, cands in (("ir_min_gain", (2, 8, 32, 128)), ("ir_fp32_cycle", (1, 2, 4, 8)),
                            ("gmres_check", (1, 2, 4))):
            res = {}
            for v in cands:
                set_knob(name, v)
                solve()
                res[v] = _wall(solve, 2)
            best = min(res, key=res.get)
            set_knob(name, best)
            out["picks"][name] = best
            out["timings_ms"][name] = res
    _CACHE[key] = out
    _save_file(key, out)
    return out
