"""Loader for libxtfem_hip.so (the gfx950 HIP build of the space-time solve path).

There is no CPU fallback: if the library is missing or fails to load, every call
raises.  ``build()`` in __graft_entry__ (or ``python -m xtfem_amd.build``) compiles it.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libxtfem_hip.so")

_LIB = None

vp = C.c_void_p
ci = C.c_int
cd = C.c_double
pi = C.POINTER(C.c_int)
pd = C.POINTER(C.c_double)


class XtError(RuntimeError):
    pass


# name -> (restype, argtypes); pointers are passed as c_void_p (numpy .ctypes.data or
# torch .data_ptr()) to keep one convention for host and device buffers.
_SIGS = {
    "xt_last_error": (C.c_char_p, []),
    "xt_kronspmv": (ci, [vp, vp, ci, vp, vp, vp, vp, vp, vp, ci, vp, vp]),
    "xt_kronspgemv": (ci, [vp, vp, cd, cd, ci, vp, vp, vp, vp, vp, vp, ci, vp, vp]),
    "xt_amux": (ci, [ci, vp, vp, vp, vp, vp]),
    "xt_ilut": (ci, [ci, vp, vp, vp, ci, cd, vp, vp, vp, ci, vp, vp, vp]),
    "xt_lusol": (ci, [ci, vp, vp, vp, vp, vp]),
    "xt_kronlusol": (ci, [vp, ci, vp, vp, vp, ci, vp, vp]),
    "xt_kronpgmres": (ci, [vp, vp, ci, vp, vp, vp, vp, vp, vp, ci, vp, vp, vp, vp, vp, ci, cd,
                           ci, vp, ci, vp, vp, vp]),
    "xt_mpipgmres": (ci, [vp, vp, ci, vp, vp, vp, vp, vp, vp, ci, vp, vp, vp, ci, cd, ci, vp,
                          ci, vp, vp, vp]),
    "xt_kokkos_host": (ci, [ci] * 9 + [vp] * 17),
    "xt_chkelesta": (ci, [ci, ci, cd, vp, vp, vp, vp, vp]),
    "xt_delelement": (ci, [ci] * 7 + [vp] * 7 + [ci, vp, vp, vp, vp, ci, vp]),
    "xt_chkcnsta": (ci, [ci, ci, ci, vp, ci, vp, vp]),
    "xt_chkndsta": (ci, [ci, ci, ci, vp, ci, vp, vp, vp, vp]),
    "xt_adddiagonal": (ci, [ci, ci, ci, vp, vp, vp, vp, ci, vp]),
    "xt_csrapplybc": (ci, [vp, vp, vp, ci, vp, vp]),
    "xt_elestatus_device": (ci, [ci, ci, cd, vp, vp, vp, vp, vp]),
    "xt_damage_zero_failed": (ci, [ci, vp, vp, vp, vp]),
    "xt_op_create": (ci, [ci, ci, ci, vp, vp, vp, vp]),
    "xt_op_destroy": (None, [vp]),
    "xt_op_nnz": (C.c_longlong, [vp, ci]),
    "xt_op_set_format": (ci, [vp, ci]),
    "xt_op_format": (ci, [vp]),
    "xt_op_blocks": (C.c_longlong, [vp, ci]),
    "xt_op_apply": (ci, [vp, ci, vp, cd, cd, vp, vp, ci, vp]),
    "xt_op_set_values": (ci, [vp, vp, vp]),
    "xt_prec_refactor_band": (ci, [vp, vp, vp, vp, vp]),
    "xt_prec_refresh_band": (ci, [vp, vp, vp, vp, ci, vp]),
    "xt_asm_create": (ci, [ci, ci, vp, vp, vp, vp]),
    "xt_asm_destroy": (None, [vp]),
    "xt_asm_pattern": (ci, [vp, ci, vp, vp, vp]),
    "xt_asm_run": (ci, [vp, vp, vp, cd, vp]),
    "xt_asm_values": (ci, [vp, ci, vp]),
    "xt_asm_copy": (ci, [vp, ci, vp, vp]),
    "xt_asm_device_pattern": (ci, [vp, ci, vp, vp]),
    "xt_asm_thermal": (ci, [vp, vp, cd, cd, cd, vp]),
    "xt_asm_thermo_load": (ci, [vp, vp, C.c_longlong, C.c_longlong, cd, cd, cd, cd, vp, vp, vp,
                                vp]),
    "xt_prec_create": (ci, [ci, vp, vp, vp, vp]),
    "xt_prec_create_band": (ci, [ci, vp, vp, vp, ci, vp]),
    "xt_prec_create_band_shifted": (ci, [ci, vp, vp, vp, vp, vp, vp, ci, vp, ci, ci, vp]),
    "xt_prec_shifts": (ci, [vp, vp, vp]),
    "xt_prec_create_band_seg": (ci, [ci, vp, vp, vp, ci, vp, ci, ci, vp]),
    "xt_prec_band_groups": (ci, [vp, vp, vp, vp]),
    "xt_slab_stats": (ci, [vp, vp, vp]),
    "xt_csr_values_in_pattern": (ci, [ci, vp, vp, vp, vp, vp, vp]),
    "xt_prec_band_info": (ci, [vp, vp, vp, vp]),
    "xt_prec_destroy": (None, [vp]),
    "xt_prec_levels": (ci, [vp, vp, vp]),
    "xt_prec_apply": (ci, [vp, ci, ci, vp, vp, vp, ci, vp]),
    "xt_prec_apply_noalloc": (ci, [vp, ci, ci, vp, vp, vp, ci, vp]),
    "xt_to_interleaved": (ci, [ci, ci, ci, vp, vp, vp]),
    "xt_from_interleaved": (ci, [ci, ci, ci, vp, vp, vp]),
    "xt_slab_solve": (ci, [vp, vp, ci, vp, vp, vp, ci, cd, ci, ci, vp, vp, vp, vp]),
    "xt_slab_solve_tol": (ci, [vp, vp, ci, vp, vp, vp, ci, cd, cd, ci, ci, vp, vp, vp, vp]),
    "xt_slab_solve_seg": (ci, [vp, vp, ci, vp, ci, vp, vp, vp, ci, cd, cd, ci, ci, vp, vp, vp,
                               vp]),
    "xt_damage_device": (ci, [ci, ci, ci] + [vp] * 18),
    "xt_damage_device_params": (ci, [ci, ci, ci] + [vp] * 19),
    "xt_slab_rhs": (ci, [ci, ci, ci, vp, vp, C.c_longlong, vp, vp, vp]),
    "xt_prof_enable": (ci, [C.c_uint]),
    "xt_graphs_enable": (ci, [ci]),
    "xt_tune_set": (ci, [ci, ci]),
    "xt_tune_get": (ci, [ci]),
    "xt_prof_read": (ci, [ci, vp, vp, vp, vp]),
    "xt_prof_read_alg": (ci, [ci, vp, vp]),
    "xt_device_sync": (ci, []),
}

EXPORTS = tuple(_SIGS)


def lib():
    """Load the HIP library (raises XtError if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise XtError(f"{LIBPATH} is not built; run `python -m xtfem_amd.build`")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = lib().xt_last_error()
        raise XtError(f"{what} failed with code {rc}: {msg.decode() if msg else ''}")
