"""Device-resident handles over the C ABI (operators, ILUT preconditioners, slab solves,
damage state).  torch supplies device memory and the current HIP stream; all compute
is in libxtfem_hip.so.  Vectors on the device use the interleaved layout
X[dof][r*q + j] (r = load case, j = temporal function)."""
from __future__ import annotations

import ctypes as C
import numpy as np
import torch

from ._lib import check, lib
from . import api


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev():
    if not torch.cuda.is_available():
        raise api.XtError("no HIP device visible: the space-time solve runs on the GPU only")
    return torch.device("cuda", torch.cuda.current_device())


def _colmajor(Ts):
    """[..., q, q] -> flat column-major per matrix (Fortran storage)."""
    Ts = np.asarray(Ts, dtype=np.float64)
    return np.ascontiguousarray(np.swapaxes(Ts, -1, -2)).reshape(Ts.shape[:-2] + (-1,))


class KronOperator:
    """Y = sum_k S_k X T_k^T on the device (kronspmv/kronspgemv, kronecker.f90:19-127)."""

    def __init__(self, mats, q=6):
        self.q = q
        self.nterms = len(mats)
        self.p = len(mats[0][0]) - 1
        keep = [(np.ascontiguousarray(ia, dtype=np.int32), np.ascontiguousarray(ja, dtype=np.int32),
                 np.ascontiguousarray(a, dtype=np.float64)) for ia, ja, a in mats]
        rp = (C.c_void_p * self.nterms)(*[k[0].ctypes.data for k in keep])
        cj = (C.c_void_p * self.nterms)(*[k[1].ctypes.data for k in keep])
        vv = (C.c_void_p * self.nterms)(*[k[2].ctypes.data for k in keep])
        h = C.c_void_p()
        _dev()
        check(lib().xt_op_create(self.p, q, self.nterms, rp, cj, vv, C.addressof(h)), "op_create")
        self.h = h
        self.nnz = [int(lib().xt_op_nnz(h, k)) for k in range(self.nterms)]
        self.blocks = int(lib().xt_op_blocks(h, 0))
        self.blocks_real = int(lib().xt_op_blocks(h, 1))

    @property
    def format(self):
        """'blocks' (node-block slices) or 'csr'."""
        return "blocks" if lib().xt_op_format(self.h) == 0 else "csr"

    @format.setter
    def format(self, name):
        check(lib().xt_op_set_format(self.h, 0 if name == "blocks" else 1), "op_set_format")

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and lib is not None:   # (module teardown clears lib)
            lib().xt_op_destroy(h)
            self.h = None

    def set_values(self, *vals):
        """New values (device FP64 tensors, one per term, in the creation pattern)."""
        ptrs = (C.c_void_p * self.nterms)(*[v.data_ptr() for v in vals])
        check(lib().xt_op_set_values(self.h, ptrs, _stream()), "op_set_values")

    def algorithmic_bytes(self, R, fp32, beta=False, fmt=None, padded=True):
        """Bytes one application moves in the stored format, each read once and written
        once: matrix (values + indices + row/slice pointers), X read, Y written.
        CSR: nnz_k (s_v + 4) + 4 (p + 1) per term.  Node blocks: per stored block 12
        values (3x3 K + diagonal M) + one int32 column node, plus slice tables."""
        sv = 4 if fp32 else 8
        n = self.p * self.q * R
        fmt = fmt or self.format
        if fmt == "blocks":
            nb = self.blocks if padded else self.blocks_real
            nsl = -(-(self.p // 3) // 64)
            mat = nb * (12 * sv + 4) + nsl * (64 * 4 + 8)
        else:
            mat = sum(nz * (sv + 4) + 4 * (self.p + 1) for nz in self.nnz)
        return mat + 2 * n * sv + (n * sv if beta else 0)

    def flops(self, R):
        c = self.q * R
        return 2 * sum(self.nnz) * c + 2 * self.nterms * self.q * self.p * c

    def tmat(self, Ts, fp32):
        """Per-load-case temporal matrices [R][nterms][q][q] (row-major) on the device."""
        t = torch.as_tensor(np.ascontiguousarray(Ts, dtype=np.float64), device=_dev())
        return t.float().contiguous() if fp32 else t.contiguous()

    def apply(self, tmat, x, y, R, alpha=1.0, beta=0.0, fp32=False):
        check(lib().xt_op_apply(self.h, R, tmat.data_ptr(), float(alpha), float(beta),
                                x.data_ptr(), y.data_ptr(), int(fp32), _stream()), "op_apply")

    def apply_host(self, A, B, X, fp32=False):
        """Y_r = (A_r (x) K + B_r (x) M) X_r for host temporal-major X [R, q*p]."""
        A = np.asarray(A).reshape(-1, self.q, self.q)
        B = np.asarray(B).reshape(-1, self.q, self.q)
        X = np.atleast_2d(X)
        R = X.shape[0]
        xi = to_interleaved(X, self.p, self.q)
        T = self.tmat(np.stack([A, B], axis=1), fp32)
        dt = torch.float32 if fp32 else torch.float64
        xin = xi.to(dt).contiguous()
        y = torch.empty_like(xin)
        self.apply(T, xin, y, R, fp32=fp32)
        return from_interleaved(y.to(torch.float64).contiguous(), self.p, self.q, R)


def to_interleaved(X, p, q):
    """host temporal-major [R, q*p] -> device interleaved FP64 [p*q*R]."""
    X = np.ascontiguousarray(np.atleast_2d(X), dtype=np.float64)
    R = X.shape[0]
    src = torch.as_tensor(X, device=_dev()).contiguous()
    dst = torch.empty(p * q * R, dtype=torch.float64, device=src.device)
    check(lib().xt_to_interleaved(p, q, R, src.data_ptr(), dst.data_ptr(), _stream()),
          "to_interleaved")
    return dst


def from_interleaved(y, p, q, R):
    dst = torch.empty(p * q * R, dtype=torch.float64, device=y.device)
    check(lib().xt_from_interleaved(p, q, R, y.data_ptr(), dst.data_ptr(), _stream()),
          "from_interleaved")
    return dst.cpu().numpy().reshape(R, q * p)


class IlutPreconditioner:
    """ILUT factors of K (ilut.f:152-495, host) with the level-scheduled GPU solves of
    kronlusol/lusol (kronecker.f90:130-178, ilut.f:2383-2432)."""

    def __init__(self, K, lfil=20, droptol=1e-4):
        alu, jlu, ju, ierr = api.ilut(K, lfil, droptol)
        if ierr != 0:
            raise api.XtError(f"ilut failed with ierr={ierr}")
        self.p = len(ju)
        self.factors = (alu, jlu, ju)
        self.nnz = int(jlu[self.p]) - 1 - (self.p + 1)
        h = C.c_void_p()
        _dev()
        check(lib().xt_prec_create(self.p, alu.ctypes.data, jlu.ctypes.data, ju.ctypes.data,
                                   C.addressof(h)), "prec_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and lib is not None:   # (module teardown clears lib)
            lib().xt_prec_destroy(h)
            self.h = None

    def levels(self):
        lo, up = C.c_int(0), C.c_int(0)
        check(lib().xt_prec_levels(self.h, C.addressof(lo), C.addressof(up)), "prec_levels")
        return lo.value, up.value

    def apply(self, minv, x, y, R, q=6, fp32=False):
        check(lib().xt_prec_apply(self.h, R, q, minv.data_ptr() if minv is not None else None,
                                  x.data_ptr(), y.data_ptr(), int(fp32), _stream()), "prec_apply")


class BandPreconditioner(IlutPreconditioner):
    """Exact factor K^-1 of a banded SPD K (the reference's live MUMPS preconditioner,
    mumps.f90:22-145 / mpilusol kronecker_mumps.f90:120-183; ilut's lfil >= n,
    droptol = 0 limit, ilut.f:2169-2170).  Block Cholesky in FP64 on the device,
    applied by dense block sweeps in FP64 or FP32."""

    def __init__(self, K, block=0, M=None, shifts=(), shifted_fp32_only=False, fp32_only=False,
                 seg=None):
        """shifts (with M): also factor K + s M for each s; slab solves then split each
        load case's dominant temporal mode onto the nearest shift (xt_prec_create_band_shifted;
        shifted_fp32_only keeps those factors in FP32 only, for the FP32 applies;
        fp32_only keeps the factor of K in FP32 only as well).  seg (DOF offsets [nseg+1]
        of independent systems of a block-diagonal K, as block_diagonal returns them):
        a banded factor per run of systems with the run's own block size
        (xt_prec_create_band_seg; no shifts)."""
        ia, ja, a = (np.ascontiguousarray(K[0], dtype=np.int32),
                     np.ascontiguousarray(K[1], dtype=np.int32),
                     np.ascontiguousarray(K[2], dtype=np.float64))
        self.p = len(ia) - 1
        self.factors = None
        self.nnz = None
        self.shifts = np.asarray(shifts, dtype=np.float64).ravel()
        h = C.c_void_p()
        _dev()
        if seg is not None:
            if len(self.shifts):
                raise ValueError("a grouped banded factor takes no shifts")
            sd = np.ascontiguousarray(seg, dtype=np.int32)
            check(lib().xt_prec_create_band_seg(
                self.p, ia.ctypes.data, ja.ctypes.data, a.ctypes.data, len(sd) - 1,
                sd.ctypes.data, int(block), 2 if fp32_only else 0, C.addressof(h)),
                "prec_create_band_seg")
        elif len(self.shifts) or fp32_only:
            if len(self.shifts):
                mi, mj, mv = (np.ascontiguousarray(M[0], dtype=np.int32),
                              np.ascontiguousarray(M[1], dtype=np.int32),
                              np.ascontiguousarray(M[2], dtype=np.float64))
                mp = (mi.ctypes.data, mj.ctypes.data, mv.ctypes.data)
            else:
                mp = (None, None, None)
            flags = (1 if shifted_fp32_only else 0) | (2 if fp32_only else 0)
            check(lib().xt_prec_create_band_shifted(
                self.p, ia.ctypes.data, ja.ctypes.data, a.ctypes.data, *mp, len(self.shifts),
                self.shifts.ctypes.data if len(self.shifts) else None, int(block), flags,
                C.addressof(h)), "prec_create_band_shifted")
        else:
            check(lib().xt_prec_create_band(self.p, ia.ctypes.data, ja.ctypes.data,
                                            a.ctypes.data, int(block), C.addressof(h)),
                  "prec_create_band")
        self.h = h
        b, nb, bw = C.c_int(0), C.c_int(0), C.c_int(0)
        check(lib().xt_prec_band_info(h, C.addressof(b), C.addressof(nb), C.addressof(bw)),
              "prec_band_info")
        self.block, self.nblocks, self.bandwidth = b.value, nb.value, bw.value
        ng = C.c_int(0)
        check(lib().xt_prec_band_groups(h, C.addressof(ng), None, None), "prec_band_groups")
        goff = np.zeros(ng.value + 1, dtype=np.int32)
        gblock = np.zeros(ng.value, dtype=np.int32)
        check(lib().xt_prec_band_groups(h, C.addressof(ng), goff.ctypes.data,
                                        gblock.ctypes.data), "prec_band_groups")
        self.group_offsets, self.group_blocks = goff, gblock

    def refactor(self, rowptr, colidx, vals, first_row=0):
        """Refactor from K on the device (1-based CSR int32/FP64 tensors, same band).
        first_row > 0: K changed only in rows and columns >= first_row since the last
        factorization; the blocks before it are kept (xt_prec_refresh_band)."""
        if first_row > 0:
            check(lib().xt_prec_refresh_band(self.h, rowptr.data_ptr(), colidx.data_ptr(),
                                             vals.data_ptr(), int(first_row), _stream()),
                  "prec_refresh_band")
            return
        check(lib().xt_prec_refactor_band(self.h, rowptr.data_ptr(), colidx.data_ptr(),
                                          vals.data_ptr(), _stream()), "prec_refactor_band")

    def matrix_bytes(self, fp32):
        """Bytes of the four block-matrix sets one apply streams (D, D^T halves counted
        as stored: whole b x b blocks; summed over the runs of a grouped factor)."""
        sv = 4 if fp32 else 8
        tot = 0
        for g, b in enumerate(self.group_blocks):
            nb = -(-int(self.group_offsets[g + 1] - self.group_offsets[g]) // int(b))
            tot += sv * int(b) * int(b) * (2 * nb + 2 * (nb - 1))
        return tot


def dominant_temporal_mode(A, B):
    """Dominant eigenvalue of A_r^-1 B_r per load case ([..., q, q] arrays; NaN where it is
    not real and positive), as the slab solver's split computes it."""
    C_ = np.linalg.solve(np.asarray(A, dtype=np.float64), np.asarray(B, dtype=np.float64))
    w = np.linalg.eigvals(C_)
    i = np.argmax(np.abs(w), axis=-1)
    lam = np.take_along_axis(w, i[..., None], axis=-1)[..., 0]
    return np.where((np.abs(lam.imag) <= 1e-10 * np.abs(lam)) & (lam.real > 0), lam.real, np.nan)


def lowest_mode(pr, M, fixed, iters=40, seed=0):
    """Estimate of the smallest eigenvalue mu_min of K phi = mu M phi: power iteration on
    K^-1 M with the device factor pr (K^-1) and a CSR operator for M.  The fixed DOFs of
    the 1-0 method (unit diagonal in K and M, per-DOF mask `fixed`) would give mu = 1:
    the iterate is kept zero there."""
    opM = KronOperator([M], q=1)
    free = torch.as_tensor(np.asarray(fixed) == 0, device=_dev())
    p = pr.p
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = (torch.rand(p, generator=g, dtype=torch.float64) - 0.5).to(_dev()) * free
    tm = torch.ones(1, dtype=torch.float64, device=_dev())
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    rho = 0.0
    for _ in range(iters):
        x /= torch.linalg.norm(x)
        opM.apply(tm, x, y, 1)
        pr.apply(None, y, z, 1, q=1)
        z *= free
        rho = float(torch.dot(x, z))
        x, z = z, x
    return 1.0 / rho


def shifts_for_cases(lam, mu_min, ratio=4.0, floor=0.05):
    """Shifts for xt_prec_create_band_shifted covering the dominant temporal eigenvalues
    lam of the load cases: cases with lam below floor * mu_min need none (the mass term
    is small against K there); the rest are grouped so that every case lies within a
    factor sqrt(ratio) of its group's shift (geometric centres)."""
    lam = np.asarray(lam, dtype=np.float64)
    big = np.sort(lam[np.isfinite(lam) & (lam >= floor * mu_min)])
    out = []
    i = 0
    while i < len(big):
        j = i
        while j + 1 < len(big) and big[j + 1] <= big[i] * ratio:
            j += 1
        out.append(float(np.sqrt(big[i] * big[j])))
        i = j + 1
    return out


class Assembly:
    """Spatial assembly on the device (mpiglbmtx_m + tricsr + csrapplybc,
    mpisubs.f90:69-139, spmtx.f90:268-407): the global patterns and each entry's element
    contributions are built once per mesh; ``run`` recomputes every element matrix for
    per-Gauss-point E and nu and sums them in the reference's order, with the fixed DOFs'
    rows and columns set by the 1-0 method.  ``K``/``M`` are device tensors of values
    in the host patterns ``K_pattern``/``M_pattern`` (1-based CSR)."""

    def __init__(self, mesh, fixed):
        dev = _dev()
        self.nn, self.ne = mesh.nn, mesh.ne
        self.p = 3 * mesh.nn
        xcoord = np.ascontiguousarray(mesh.xcoord, dtype=np.float64)
        xconn = np.ascontiguousarray(mesh.xconn, dtype=np.int32)
        fx = np.ascontiguousarray(np.asarray(fixed) != 0, dtype=np.int32)
        h = C.c_void_p()
        check(lib().xt_asm_create(self.nn, self.ne, xcoord.ctypes.data, xconn.ctypes.data,
                                  fx.ctypes.data, C.addressof(h)), "asm_create")
        self.h = h
        self.K_pattern = self._pattern(0)
        self.M_pattern = self._pattern(1)
        t = lambda a: torch.as_tensor(a, device=dev)
        self.K_rowptr, self.K_col = t(self.K_pattern[0]), t(self.K_pattern[1])
        self.K = torch.zeros(len(self.K_pattern[1]), dtype=torch.float64, device=dev)
        self.M = torch.zeros(len(self.M_pattern[1]), dtype=torch.float64, device=dev)
        self.Nxyz = [torch.zeros(64 * self.ne, dtype=torch.float64, device=dev) for _ in range(3)]

    def _pattern(self, which):
        nnz = C.c_longlong(0)
        check(lib().xt_asm_pattern(self.h, which, C.addressof(nnz), None, None), "asm_pattern")
        ia = np.zeros((self.nn if which == 2 else self.p) + 1, dtype=np.int32)
        ja = np.zeros(nnz.value, dtype=np.int32)
        check(lib().xt_asm_pattern(self.h, which, C.addressof(nnz), ia.ctypes.data,
                                   ja.ctypes.data), "asm_pattern")
        return ia, ja

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and lib is not None:   # (module teardown clears lib)
            lib().xt_asm_destroy(h)
            self.h = None

    def run(self, ym, nu, rho):
        """Assemble for E, nu (scalars, or per Gauss point [8*ne], host or device)."""
        dev = self.K.device
        g = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).expand(8 * self.ne).contiguous()
        ymd, nud = g(ym), g(nu)
        s = _stream()
        check(lib().xt_asm_run(self.h, ymd.data_ptr(), nud.data_ptr(), float(rho), s), "asm_run")
        for w, dst in ((0, self.K), (1, self.M), (2, self.Nxyz[0]), (3, self.Nxyz[1]),
                       (4, self.Nxyz[2])):
            check(lib().xt_asm_copy(self.h, w, dst.data_ptr(), s), "asm_copy")
        return self.K, self.M

    def host(self):
        """(K, M) as host CSR triples (1-based), values of the last ``run``."""
        return ((self.K_pattern[0], self.K_pattern[1], self.K.cpu().numpy()),
                (self.M_pattern[0], self.M_pattern[1], self.M.cpu().numpy()))

    def thermal(self, tfixed, kappa, cp, rho):
        """Thermal conductivity and capacity (elemtx_t + mpiglbmtx_t, elesubs.f90:60-116,
        mpisubs.f90:141-188) on the device, with the 1-0 method on the nodes where
        ``tfixed`` [nn] is nonzero.  Sets ``T_pattern`` (1-based node CSR shared by both)
        and the device value tensors ``Kt``, ``Mt`` (with BCs) and ``Kt0``, ``Mt0``
        (without).  Returns (Kt, Mt)."""
        tf = np.ascontiguousarray(np.asarray(tfixed) != 0, dtype=np.int32)
        if tf.shape != (self.nn,):
            raise api.XtError(f"tfixed must have one entry per node ({self.nn})")
        s = _stream()
        check(lib().xt_asm_thermal(self.h, tf.ctypes.data, float(kappa), float(cp), float(rho), s),
              "asm_thermal")
        if not hasattr(self, "T_pattern"):
            self.T_pattern = self._pattern(2)
            dev = self.K.device
            self.T_rowptr = torch.as_tensor(self.T_pattern[0], device=dev)
            self.T_col = torch.as_tensor(self.T_pattern[1], device=dev)
            z = lambda: torch.zeros(len(self.T_pattern[1]), dtype=torch.float64, device=dev)
            self.Kt, self.Mt, self.Kt0, self.Mt0 = z(), z(), z(), z()
        for w, dst in ((5, self.Kt), (6, self.Mt), (7, self.Kt0), (8, self.Mt0)):
            check(lib().xt_asm_copy(self.h, w, dst.data_ptr(), s), "asm_copy")
        return self.Kt, self.Mt

    def thermo_load(self, T, tstride, toff, Tref, alpha, ym0, nu0):
        """material_update and the thermal load vv of one slab (thermoelastic.f90:1-78,
        elesubs.f90:119-167) from the device temperatures T[n*tstride + toff].
        Returns device tensors (ym [8*ne], nu [8*ne], vv [3*nn]), reused across calls."""
        if T.dtype != torch.float64 or not T.is_contiguous() or T.device != self.K.device:
            raise api.XtError("thermo_load: T must be a contiguous FP64 tensor on the device")
        if not (0 <= toff < tstride) or T.numel() < self.nn * tstride:
            raise api.XtError("thermo_load: T too short for nn * tstride")
        if not hasattr(self, "ym_t"):
            dev = self.K.device
            self.ym_t = torch.zeros(8 * self.ne, dtype=torch.float64, device=dev)
            self.nu_t = torch.zeros(8 * self.ne, dtype=torch.float64, device=dev)
            self.vv = torch.zeros(self.p, dtype=torch.float64, device=dev)
        check(lib().xt_asm_thermo_load(self.h, T.data_ptr(), int(tstride), int(toff), float(Tref),
                                       float(alpha), float(ym0), float(nu0), self.ym_t.data_ptr(),
                                       self.nu_t.data_ptr(), self.vv.data_ptr(), _stream()),
              "asm_thermo_load")
        return self.ym_t, self.nu_t, self.vv


def slab_solve(op, pr, A, B, rhs, x0=None, im=30, eps=1e-8, maxits=500, mode=1, xtol=0.0):
    """Solve (A_r (x) K + B_r (x) M) x_r = rhs_r for every load case r on the GPU.

    mode 0: FP64 GMRES(im) with the FP64 Kronecker preconditioner.
    mode 1: GMRES iterative refinement: FP64 true residual, inner flexible GMRES cycles
            with the FP32 operator and FP32 preconditioner sweeps.
    mode 2: FP64 flexible GMRES(im) with the FP32 preconditioner sweeps.
    xtol (mode 1): also require the last refinement correction to be at most xtol times
    the solution (forward-error test, xt_slab_solve_tol).
    rhs: host [R, q*p] temporal-major.  Returns (x [R, q*p], its, ierr, relres)."""
    A = np.asarray(A, dtype=np.float64).reshape(-1, op.q, op.q)
    B = np.asarray(B, dtype=np.float64).reshape(-1, op.q, op.q)
    rhs = np.atleast_2d(rhs)
    R = rhs.shape[0]
    b = to_interleaved(rhs, op.p, op.q)
    x = torch.zeros_like(b) if x0 is None else to_interleaved(x0, op.p, op.q)
    x, its, ierr, relres = slab_solve_device(op, pr, A, B, b, x, R, im, eps, maxits, mode,
                                             xtol)
    return from_interleaved(x, op.p, op.q, R), its, ierr, relres


def slab_solve_device(op, pr, A, B, b, x, R, im=30, eps=1e-8, maxits=500, mode=1, xtol=0.0):
    """Device-vector form of slab_solve (b, x interleaved FP64 device tensors; x in/out)."""
    ab = np.ascontiguousarray(np.stack([_colmajor(A), _colmajor(B)], axis=1))
    its = np.zeros(R, dtype=np.int32)
    ierr = np.zeros(R, dtype=np.int32)
    relres = np.zeros(R)
    if xtol > 0.0:
        check(lib().xt_slab_solve_tol(op.h, pr.h, R, ab.ctypes.data, b.data_ptr(), x.data_ptr(),
                                      int(im), float(eps), float(xtol), int(maxits), int(mode),
                                      its.ctypes.data, ierr.ctypes.data, relres.ctypes.data,
                                      _stream()), "slab_solve_tol")
    else:
        check(lib().xt_slab_solve(op.h, pr.h, R, ab.ctypes.data, b.data_ptr(), x.data_ptr(),
                                  int(im), float(eps), int(maxits), int(mode), its.ctypes.data,
                                  ierr.ctypes.data, relres.ctypes.data, _stream()), "slab_solve")
    return x, its, ierr, relres


def slab_stats(op):
    """(it32, it64): iterations of the last slab solve on op whose preconditioner sweeps
    ran in FP32 and in FP64, summed over its instances (xt_slab_stats)."""
    a, b = C.c_longlong(), C.c_longlong()
    check(lib().xt_slab_stats(op.h, C.addressof(a), C.addressof(b)), "slab_stats")
    return a.value, b.value


def block_diagonal(mats):
    """Stack 1-based CSR matrices (ia, ja, a) into one block-diagonal CSR; returns it and the
    row offsets of the blocks [n_blocks + 1]."""
    ia_out, ja_out, a_out = [np.zeros(1, dtype=np.int64)], [], []
    off = [0]
    nnz = 0
    for ia, ja, a in mats:
        ia = np.asarray(ia, dtype=np.int64)
        n = len(ia) - 1
        ia_out.append(ia[1:] - 1 + nnz)
        ja_out.append(np.asarray(ja, dtype=np.int64) + off[-1])
        a_out.append(np.asarray(a, dtype=np.float64))
        nnz += int(ia[-1] - 1)
        off.append(off[-1] + n)
    ia = (np.concatenate(ia_out) + 1).astype(np.int32)
    return (ia, np.concatenate(ja_out).astype(np.int32), np.concatenate(a_out)), np.array(off)


def slab_solve_seg(op, pr, seg_dof, A, B, b, x, R, im=30, eps=1e-8, maxits=500, mode=1,
                   xtol=0.0):
    """xt_slab_solve_seg: independent systems stacked block-diagonally (rows
    seg_dof[s]..seg_dof[s+1]-1), R load cases each, one instance per (system, case).
    b, x: interleaved FP64 device tensors; returns (x, its, ierr, relres) with the
    per-instance arrays shaped [nseg, R]."""
    nseg = len(seg_dof) - 1
    sd = np.ascontiguousarray(seg_dof, dtype=np.int64)
    ab = np.ascontiguousarray(np.stack([_colmajor(A), _colmajor(B)], axis=1))
    its = np.zeros(nseg * R, dtype=np.int32)
    ierr = np.zeros(nseg * R, dtype=np.int32)
    relres = np.zeros(nseg * R)
    check(lib().xt_slab_solve_seg(op.h, pr.h, nseg, sd.ctypes.data, R, ab.ctypes.data,
                                  b.data_ptr(), x.data_ptr(), int(im), float(eps), float(xtol),
                                  int(maxits), int(mode), its.ctypes.data, ierr.ctypes.data,
                                  relres.ctypes.data, _stream()), "slab_solve_seg")
    return x, its.reshape(nseg, R), ierr.reshape(nseg, R), relres.reshape(nseg, R)


class DamageState:
    """Per-Gauss-point damage state resident in HBM across slabs, with the C3D8 shape
    derivative tables; ``step`` runs the damage kernel (kokkos_kernel.h:44-344) for one
    slab and ``failed_elements`` applies chkelesta (delelement.f90:109-136)."""

    FIELDS6 = ("eps_old", "eps_pl", "sig_eff", "sig_back")
    FIELDS1 = ("p", "D", "ws", "del_po", "sig_d")

    def __init__(self, mesh, Nx, Ny, Nz, ym, nu, params=None):
        """params: the damage-law constants (sig_f, sig_u, eps_D, Cy, h, S, s) or None for
        the reference's (damage.h:9-17)."""
        dev = _dev()
        self.ne, self.nn = mesh.ne, mesh.nn
        self.params = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        ngp = 8 * mesh.ne
        t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), device=dev).to(dt).contiguous()
        self.nodes = t(np.repeat(mesh.xconn, 8, axis=0).ravel(), torch.int32)
        self.Nx, self.Ny, self.Nz = t(np.ravel(Nx)), t(np.ravel(Ny)), t(np.ravel(Nz))
        self.ym = t(np.broadcast_to(ym, (ngp,)))
        self.nu = t(np.broadcast_to(nu, (ngp,)))
        self.s = {k: torch.zeros(6 * ngp, dtype=torch.float64, device=dev) for k in self.FIELDS6}
        self.s.update({k: torch.zeros(ngp, dtype=torch.float64, device=dev) for k in self.FIELDS1})
        self.gpsta = torch.ones(ngp, dtype=torch.int32, device=dev)
        self.elesta = torch.ones(mesh.ne, dtype=torch.int32, device=dev)

    def step(self, sol_dev, Nt_dev, nip):
        """sol_dev: temporal-major FP64 device vector [6*3*nn]; Nt_dev: [nip*6]."""
        s = self.s
        check(lib().xt_damage_device_params(
            self.ne, self.nn, nip, self.nodes.data_ptr(), self.Nx.data_ptr(), self.Ny.data_ptr(),
            self.Nz.data_ptr(), Nt_dev.data_ptr(), sol_dev.data_ptr(), self.ym.data_ptr(),
            self.nu.data_ptr(), *[s[k].data_ptr() for k in self.FIELDS6],
            *[s[k].data_ptr() for k in self.FIELDS1],
            None if self.params is None else self.params.ctypes.data, _stream()),
            "damage_device")
        # kokkos_damage.f90:102-107: failed Gauss points carry D = ws = 0
        check(lib().xt_damage_zero_failed(8 * self.ne, self.gpsta.data_ptr(), s["D"].data_ptr(),
                                          s["ws"].data_ptr(), _stream()), "damage_zero_failed")

    def failed_elements(self, Dc):
        """chkelesta (delelement.f90:109-136) on the device state: marks the newly failed
        elements and their Gauss points in elesta / gpsta and returns their 1-based ids
        in increasing order."""
        flags = torch.empty(self.ne, dtype=torch.int32, device=self.gpsta.device)
        check(lib().xt_elestatus_device(self.ne, 8, float(Dc), self.s["D"].data_ptr(),
                                        self.elesta.data_ptr(), self.gpsta.data_ptr(),
                                        flags.data_ptr(), _stream()), "elestatus_device")
        return (np.flatnonzero(flags.cpu().numpy()) + 1).astype(np.int32)

    def host_state(self):
        return {k: v.cpu().numpy() for k, v in self.s.items()}
