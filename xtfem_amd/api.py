"""Host-side mirror of the reference subroutines on the space-time solve path.

Same names, argument meaning and error behaviour as the reference routines; every
call runs on the GPU through libxtfem_hip.so (C ABI declared in include/xtfem.h),
except ``ilut`` and the element-deletion edits of K (``delelement``, ``chkcnsta``,
``chkndsta``, ``adddiagonal``, ``csrapplybc``), which are host C++ in the same library.

Conventions (the reference's): CSR triples ``(ia, ja, a)`` 1-based int32 / float64,
temporal matrices q x q numpy arrays, space-time vectors temporal-major
``x[j*p + dof]``.
"""
from __future__ import annotations

import ctypes as C
import numpy as np

from ._lib import check, lib, XtError


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _P(a):
    return a.ctypes.data


def _fmat(A):
    """q x q -> flat column-major (Fortran storage of a(i,j))."""
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).T).ravel()


def kronspmv(A, B, K, M, x):
    """y = (A (x) K + B (x) M) x.  kronecker.f90:19-85"""
    q = A.shape[0]
    ki, kj, kv = _i32(K[0]), _i32(K[1]), _f64(K[2])
    mi, mj, mv = _i32(M[0]), _i32(M[1]), _f64(M[2])
    p = len(ki) - 1
    a, b, x = _fmat(A), _fmat(B), _f64(x)
    y = np.zeros(p * q)
    check(lib().xt_kronspmv(_P(a), _P(b), q, _P(ki), _P(kj), _P(kv), _P(mi), _P(mj), _P(mv), p,
                            _P(x), _P(y)), "kronspmv")
    return y


def kronspgemv(A, B, alpha, beta, K, M, x, y):
    """y := alpha (A (x) K + B (x) M) x + beta y.  kronecker.f90:88-127"""
    q = A.shape[0]
    ki, kj, kv = _i32(K[0]), _i32(K[1]), _f64(K[2])
    mi, mj, mv = _i32(M[0]), _i32(M[1]), _f64(M[2])
    p = len(ki) - 1
    a, b, x = _fmat(A), _fmat(B), _f64(x)
    y = _f64(y).copy()
    check(lib().xt_kronspgemv(_P(a), _P(b), float(alpha), float(beta), q, _P(ki), _P(kj), _P(kv),
                              _P(mi), _P(mj), _P(mv), p, _P(x), _P(y)), "kronspgemv")
    return y


def amux(ia, ja, a, x):
    """y = A x (CSR).  ilut.f:3-44"""
    ia, ja, a, x = _i32(ia), _i32(ja), _f64(a), _f64(x)
    n = len(ia) - 1
    y = np.zeros(n)
    check(lib().xt_amux(n, _P(x), _P(y), _P(a), _P(ja), _P(ia)), "amux")
    return y


def ilut(K, lfil, droptol, iwk=None):
    """Dual-threshold ILU -> (alu, jlu, ju, ierr), 1-based MSR.  ilut.f:152-495"""
    ia, ja, a = _i32(K[0]), _i32(K[1]), _f64(K[2])
    n = len(ia) - 1
    if iwk is None:
        iwk = n + 2 + n * (2 * lfil + 1) + 16
    alu = np.zeros(iwk)
    jlu = np.zeros(iwk, dtype=np.int32)
    ju = np.zeros(n, dtype=np.int32)
    w = np.zeros(n + 1)
    jw = np.zeros(2 * n, dtype=np.int32)
    ierr = C.c_int(0)
    check(lib().xt_ilut(n, _P(a), _P(ja), _P(ia), int(lfil), float(droptol), _P(alu), _P(jlu),
                        _P(ju), int(iwk), _P(w), _P(jw), C.addressof(ierr)), "ilut")
    nnz = int(jlu[n]) - 1 if ierr.value == 0 else 0
    return alu[:max(nnz, n + 1)].copy(), jlu[:max(nnz, n + 1)].copy(), ju, ierr.value


def lusol(alu, jlu, ju, y):
    """(LU) x = y with ilut's factors.  ilut.f:2383-2432"""
    alu, jlu, ju, y = _f64(alu), _i32(jlu), _i32(ju), _f64(y)
    n = len(ju)
    x = np.zeros(n)
    check(lib().xt_lusol(n, _P(y), _P(x), _P(alu), _P(jlu), _P(ju)), "lusol")
    return x


def kronlusol(A, alu, jlu, ju, x):
    """Solve (A (x) LU) y = x.  kronecker.f90:130-178"""
    q = A.shape[0]
    alu, jlu, ju, x = _f64(alu), _i32(jlu), _i32(ju), _f64(x)
    p = len(ju)
    a = _fmat(A)
    y = np.zeros(p * q)
    rc = lib().xt_kronlusol(_P(a), q, _P(alu), _P(jlu), _P(ju), p, _P(x), _P(y))
    if rc == -3:
        raise np.linalg.LinAlgError("Matrix is numerically singular!")
    check(rc, "kronlusol")
    return y


def kronpgmres(A, B, K, M, prec, rhs, sol0=None, im=30, eps=1e-8, maxits=500):
    """Right-preconditioned restarted GMRES with the Kronecker ILUT preconditioner.
    kronecker.f90:181-365.  Returns (sol, ierr, its, hist) with hist the residual
    history ro/ro0 that the reference writes to unit iout."""
    q = A.shape[0]
    ki, kj, kv = _i32(K[0]), _i32(K[1]), _f64(K[2])
    mi, mj, mv = _i32(M[0]), _i32(M[1]), _f64(M[2])
    p = len(ki) - 1
    alu, jlu, ju = _f64(prec[0]), _i32(prec[1]), _i32(prec[2])
    a, b = _fmat(A), _fmat(B)
    rhs = _f64(rhs).copy()
    sol = np.zeros(p * q) if sol0 is None else _f64(sol0).copy()
    cap = maxits + 64
    hist = np.zeros(cap)
    nh, its, ierr = C.c_int(0), C.c_int(0), C.c_int(0)
    check(lib().xt_kronpgmres(_P(a), _P(b), q, _P(ki), _P(kj), _P(kv), _P(mi), _P(mj), _P(mv), p,
                              _P(alu), _P(jlu), _P(ju), _P(rhs), _P(sol), int(im), float(eps),
                              int(maxits), _P(hist), cap, C.addressof(nh), C.addressof(its),
                              C.addressof(ierr)), "kronpgmres")
    return sol, ierr.value, its.value, hist[:nh.value].copy()


def mpipgmres(A, B, K, M, pr, rhs, sol0=None, im=30, eps=1e-8, maxits=500):
    """The live path's GMRES, kronecker_mumps.f90:186-415, with the preconditioner handle
    pr (e.g. device.BandPreconditioner(K), the exact factor MUMPS gives the reference) in
    place of mumps_par.  Returns (sol, ierr, its, hist) as kronpgmres."""
    q = A.shape[0]
    ki, kj, kv = _i32(K[0]), _i32(K[1]), _f64(K[2])
    mi, mj, mv = _i32(M[0]), _i32(M[1]), _f64(M[2])
    p = len(ki) - 1
    a, b = _fmat(A), _fmat(B)
    rhs = _f64(rhs).copy()
    sol = np.zeros(p * q) if sol0 is None else _f64(sol0).copy()
    cap = maxits + 64
    hist = np.zeros(cap)
    nh, its, ierr = C.c_int(0), C.c_int(0), C.c_int(0)
    check(lib().xt_mpipgmres(_P(a), _P(b), q, _P(ki), _P(kj), _P(kv), _P(mi), _P(mj), _P(mv), p,
                             pr.h, _P(rhs), _P(sol), int(im), float(eps), int(maxits), _P(hist),
                             cap, C.addressof(nh), C.addressof(its), C.addressof(ierr)),
          "mpipgmres")
    return sol, ierr.value, its.value, hist[:nh.value].copy()


def kokkos_host(ne, nn, nip, nodes, Nx, Ny, Nz, Nt, sol, ym, nu, state):
    """Damage functor over all Gauss points.  kokkos_host.cpp:37-190 (kokkos_kernel.h:44-344).
    Returns a new state dict (eps_old, eps_pl, sig_eff, sig_back, p, D, ws, del_po, sig_d)."""
    st = {k: _f64(v).copy() for k, v in state.items()}
    nodes, Nx, Ny, Nz, Nt = _i32(nodes), _f64(Nx), _f64(Ny), _f64(Nz), _f64(Nt)
    sol, ym, nu = _f64(sol), _f64(ym), _f64(nu)
    check(lib().xt_kokkos_host(0, ne, nn, 3, 8, 3, 8, 6, nip, _P(nodes), _P(Nx), _P(Ny), _P(Nz),
                               _P(Nt), _P(sol), _P(ym), _P(nu), _P(st["eps_old"]),
                               _P(st["eps_pl"]), _P(st["sig_eff"]), _P(st["sig_back"]),
                               _P(st["p"]), _P(st["D"]), _P(st["ws"]), _P(st["del_po"]),
                               _P(st["sig_d"])), "kokkos_host")
    return st


def kokkos_damage(mesh, Nx, Ny, Nz, ym, nu, w, tcoorde, sol, gpsta, nip, state):
    """kokkos_damage.f90:1-112: temporal table Nt, flattening, kernel, then D = ws = 0 at
    failed Gauss points.  Nx/Ny/Nz: [8*ne, 8] (Gauss point, local node)."""
    from .slab import damage_nt
    Nt = damage_nt(tcoorde, w, nip)
    nodes = np.repeat(_i32(mesh.xconn), 8, axis=0)  # node ids per Gauss point
    st = kokkos_host(mesh.ne, mesh.nn, nip, nodes.ravel(), Nx.ravel(), Ny.ravel(), Nz.ravel(),
                     Nt.ravel(), sol, ym, nu, state)
    failed = np.asarray(gpsta) == 0
    st["D"][failed] = 0.0
    st["ws"][failed] = 0.0
    return st


def chkelesta(ne, ngpe, Dc, D, elesta=None, gpsta=None):
    """Element status from damage.  delelement.f90:109-136 -> (nefail, elesta, gpsta)."""
    el = np.ones(ne, dtype=np.int32) if elesta is None else _i32(elesta).copy()
    gp = np.ones(ne * ngpe, dtype=np.int32) if gpsta is None else _i32(gpsta).copy()
    D = _f64(D)
    nefail = np.zeros(max(ne, 1), dtype=np.int32)
    nef = C.c_int(0)
    check(lib().xt_chkelesta(ne, ngpe, float(Dc), _P(D), _P(el), _P(gp), C.addressof(nef),
                             _P(nefail)), "chkelesta")
    return nefail[:nef.value].copy(), el, gp


def _csr_in(S, extra=0):
    """Copy of a 1-based CSR triple with room for ``extra`` more entries."""
    ia = _i32(S[0]).copy()
    nz = int(ia[-1]) - 1
    ja = np.zeros(nz + extra, dtype=np.int32)
    a = np.zeros(nz + extra)
    ja[:nz] = S[1][:nz]
    a[:nz] = S[2][:nz]
    return ia, ja, a, C.c_int(nz)


def delelement(xconn, xcoord, dns, dnxyz, sd, stiff, wqs, nefail, K):
    """K - sum of the failed elements' stiffness matrices (delelement.f90:1-59).

    xconn [ne, nnse] 1-based, xcoord [nn, 3], dns [3*ngpe, nnse],
    dnxyz [9*ngpe, 3*nnse], sd [6, 9], stiff [6, 6] (one material matrix for every
    Gauss point), wqs [nquads] as setshapefunc_m / setsdstiff give them; nefail
    1-based failed element ids.  Returns the new CSR triple."""
    xconn = np.asfortranarray(xconn, dtype=np.int32)
    xcoord = np.asfortranarray(xcoord, dtype=np.float64)
    dns = np.asfortranarray(dns, dtype=np.float64)
    dnxyz = np.asfortranarray(dnxyz, dtype=np.float64)
    sd = np.asfortranarray(sd, dtype=np.float64)
    stiff = np.asfortranarray(stiff, dtype=np.float64)
    wqs = _f64(wqs)
    nefail = _i32(nefail)
    ne, nnse = xconn.shape
    nn = xcoord.shape[0]
    ngpe = dns.shape[0] // 3
    ki, kj, kv, knz = _csr_in(K, extra=len(nefail) * (3 * nnse) ** 2)
    ndof = len(ki) - 1
    check(lib().xt_delelement(nn, ne, 3, nnse, ngpe, 6, len(wqs), _P(xconn), _P(xcoord),
                              _P(dns), _P(dnxyz), _P(sd), _P(stiff), _P(wqs), len(nefail),
                              _P(nefail), _P(ki), _P(kj), _P(kv), ndof, C.addressof(knz)),
          "delelement")
    # views of the staging arrays (a few element matrices' room beyond knz): no copy of K
    return ki, kj[:knz.value], kv[:knz.value]


def chkcnsta(nn, xconn, nefail, necnt):
    """necnt minus one per failed element at each of its nodes (delelement.f90:139-161)."""
    xconn = np.asfortranarray(xconn, dtype=np.int32)
    nefail = _i32(nefail)
    necnt = _i32(necnt).copy()
    check(lib().xt_chkcnsta(nn, xconn.shape[0], xconn.shape[1], _P(xconn), len(nefail),
                            _P(nefail), _P(necnt)), "chkcnsta")
    return necnt


def chkndsta(nn, xconn, nefail, necnt):
    """Orphan nodes of the failed elements, in discovery order (delelement.f90:164-201)."""
    xconn = np.asfortranarray(xconn, dtype=np.int32)
    nefail = _i32(nefail)
    necnt = _i32(necnt)
    ndfail = np.zeros(nn, dtype=np.int32)
    ndf = C.c_int(0)
    check(lib().xt_chkndsta(nn, xconn.shape[0], xconn.shape[1], _P(xconn), len(nefail),
                            _P(nefail), _P(necnt), C.addressof(ndf), _P(ndfail)), "chkndsta")
    return ndfail[:ndf.value].copy()


def adddiagonal(ndofn, ndfail, K):
    """K + I on the DOFs of the orphan nodes (delelement.f90:204-247); K's own arrays
    when there are none (the reference's loop then adds nothing)."""
    ndfail = _i32(ndfail)
    if len(ndfail) == 0:
        return _i32(K[0]), _i32(K[1]), np.ascontiguousarray(K[2], dtype=np.float64)
    ki, kj, kv, knz = _csr_in(K, extra=len(ndfail) * ndofn)
    check(lib().xt_adddiagonal(8, ndofn, len(ndfail), _P(ndfail), _P(ki), _P(kj), _P(kv),
                               len(ki) - 1, C.addressof(knz)), "adddiagonal")
    return ki, kj[:knz.value], kv[:knz.value]


def csrapplybc(S, bc):
    """1-0 method on the DOFs with bc != 0 (spmtx.f90:268-307)."""
    ia, ja, a, nz = _csr_in(S)
    bc = _i32(bc)
    check(lib().xt_csrapplybc(_P(ia), _P(ja), _P(a), len(ia) - 1, C.addressof(nz), _P(bc)),
          "csrapplybc")
    return ia, ja[:nz.value], a[:nz.value]


def csr_values_in_pattern(P, S):
    """Values of the 1-based CSR S in the pattern of the 1-based CSR P (zeros where S has
    no entry); XtError when S has an entry outside P (xt_csr_values_in_pattern)."""
    pia, pja = _i32(P[0]), _i32(P[1])
    ia, ja, a = _i32(S[0]), _i32(S[1]), np.ascontiguousarray(S[2], dtype=np.float64)
    out = np.zeros(len(pja))
    check(lib().xt_csr_values_in_pattern(len(pia) - 1, _P(pia), _P(pja), _P(ia), _P(ja), _P(a),
                                         _P(out)), "csr_values_in_pattern")
    return out


__all__ = ["kronspmv", "kronspgemv", "amux", "ilut", "lusol", "kronlusol", "kronpgmres",
           "kokkos_host", "kokkos_damage", "chkelesta", "delelement", "chkcnsta", "chkndsta",
           "adddiagonal", "csrapplybc", "csr_values_in_pattern", "XtError"]
