"""ctypes front end of the C oracle (``liboracle.so``, built from xtfem_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the CPU-baseline leg of bench.py, always as the checker.  The product package
(xtfem_amd) never imports it.

Conventions are the reference's: 1-based CSR (int32), column-major q x q temporal
matrices, temporal-major space-time vectors ``x[j*p + dof]``.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
# spatial solve of or_kronpgmres_cb: (ctx, p, q, x [q*p], w [q*p]) -> 0 on success
SPATIAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                         C.POINTER(C.c_double))
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        ci, cd = C.c_int, C.c_double
        L.or_amux.argtypes = [ci, _d, _d, _d, _i, _i]
        L.or_kronspmv.argtypes = [_d, _d, ci, _i, _i, _d, _i, _i, _d, ci, _d, _d]
        L.or_kronspgemv.argtypes = [_d, _d, cd, cd, ci, _i, _i, _d, _i, _i, _d, ci, _d, _d]
        L.or_ilut.argtypes = [ci, _d, _i, _i, ci, cd, _d, _i, _i, ci, _d, _i]
        L.or_ilut.restype = ci
        L.or_lusol.argtypes = [ci, _d, _d, _d, _i, _i]
        L.or_inverse.argtypes = [ci, _d, _d]
        L.or_inverse.restype = ci
        L.or_kronlusol.argtypes = [_d, ci, _d, _i, _i, ci, _d, _d]
        L.or_kronlusol.restype = ci
        L.or_kronpgmres.argtypes = [_d, _d, ci, _i, _i, _d, _i, _i, _d, ci, _d, _i, _i,
                                    _d, _d, ci, cd, ci, _d, ci, C.POINTER(ci), C.POINTER(ci)]
        L.or_kronpgmres.restype = ci
        L.or_bandchol.argtypes = [ci, ci, _d]
        L.or_bandchol.restype = ci
        L.or_bandsol.argtypes = [ci, ci, _d, _d, _d]
        L.or_kronbandsol.argtypes = [_d, ci, _d, ci, ci, _d, _d]
        L.or_kronbandsol.restype = ci
        L.or_kronpgmres_band.argtypes = [_d, _d, ci, _i, _i, _d, _i, _i, _d, ci, _d, ci,
                                         _d, _d, ci, cd, ci, _d, ci, C.POINTER(ci),
                                         C.POINTER(ci)]
        L.or_kronpgmres_band.restype = ci
        L.or_kronpgmres_cb.argtypes = [_d, _d, ci, _i, _i, _d, _i, _i, _d, ci, SPATIAL_FN,
                                       C.c_void_p, _d, _d, ci, cd, ci, _d, ci, C.POINTER(ci),
                                       C.POINTER(ci)]
        L.or_kronpgmres_cb.restype = ci
        L.or_stcoeff_m.argtypes = [cd, cd, cd, _d, _d, _d, _d, _d, _d]
        L.or_stcoeff_t.argtypes = [cd, cd, _d, _d, _d, _d, _d, _d]
        L.or_getfextc_m.argtypes = [cd, cd, cd, _d, _d]
        L.or_damage_nt.argtypes = [_d, cd, ci, _d]
        L.or_kokkos_host.argtypes = [ci] * 9 + [_i, _d, _d, _d, _d, _d, _d, _d] + [_d] * 9
        L.or_damage_step.argtypes = [_d, cd, cd, _d, _d, _d, _d, _d]
        L.or_damage_set_params.argtypes = [C.c_void_p]
        L.or_chkelesta.argtypes = [ci, ci, cd, _d, _i, _i, _i]
        L.or_chkelesta.restype = ci
        _LIB = L
    return _LIB


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _fmat(a):
    """q x q matrix -> flat column-major float64 (Fortran order)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).ravel()


def amux(ia, ja, a, x):
    """ilut.f:3-44"""
    n = len(ia) - 1
    y = np.zeros(n)
    lib().or_amux(n, _f64(x), y, _f64(a), _i32(ja), _i32(ia))
    return y


def kronspmv(A, B, K, M, x):
    """kronecker.f90:19-85. K, M are (ia, ja, a) 1-based CSR triples."""
    q = A.shape[0]
    p = len(K[0]) - 1
    y = np.zeros(p * q)
    lib().or_kronspmv(_fmat(A), _fmat(B), q, _i32(K[0]), _i32(K[1]), _f64(K[2]),
                      _i32(M[0]), _i32(M[1]), _f64(M[2]), p, _f64(x), y)
    return y


def kronspgemv(A, B, alpha, beta, K, M, x, y):
    """kronecker.f90:88-127"""
    q = A.shape[0]
    p = len(K[0]) - 1
    y = _f64(y).copy()
    lib().or_kronspgemv(_fmat(A), _fmat(B), alpha, beta, q, _i32(K[0]), _i32(K[1]), _f64(K[2]),
                        _i32(M[0]), _i32(M[1]), _f64(M[2]), p, _f64(x), y)
    return y


def ilut(K, lfil, droptol, iwk=None):
    """ilut.f:152-495. Returns (alu, jlu, ju, ierr) in 1-based MSR."""
    ia, ja, a = K
    n = len(ia) - 1
    if iwk is None:
        iwk = n + 2 + n * (2 * lfil + 1) + 16
    alu = np.zeros(iwk)
    jlu = np.zeros(iwk, dtype=np.int32)
    ju = np.zeros(n, dtype=np.int32)
    w = np.zeros(n + 1)
    jw = np.zeros(2 * n, dtype=np.int32)
    ierr = lib().or_ilut(n, _f64(a), _i32(ja), _i32(ia), lfil, droptol, alu, jlu, ju, iwk, w, jw)
    nnz = int(jlu[n]) - 1 if ierr == 0 else 0
    return alu[:max(nnz, n + 1)].copy(), jlu[:max(nnz, n + 1)].copy(), ju, ierr


def lusol(alu, jlu, ju, y):
    """ilut.f:2383-2432"""
    n = len(ju)
    x = np.zeros(n)
    lib().or_lusol(n, _f64(y), x, _f64(alu), _i32(jlu), _i32(ju))
    return x


def inverse(A):
    """kronecker.f90:368-395 (LAPACK DGETRF/DGETRI restated)."""
    q = A.shape[0]
    out = np.zeros(q * q)
    if lib().or_inverse(q, _fmat(A), out):
        raise np.linalg.LinAlgError("Matrix is numerically singular!")
    return out.reshape(q, q).T.copy()


def kronlusol(A, alu, jlu, ju, x):
    """kronecker.f90:130-178"""
    q = A.shape[0]
    p = len(ju)
    y = np.zeros(p * q)
    if lib().or_kronlusol(_fmat(A), q, _f64(alu), _i32(jlu), _i32(ju), p, _f64(x), y):
        raise np.linalg.LinAlgError("Matrix is numerically singular!")
    return y


def kronpgmres(A, B, K, M, prec, rhs, sol0=None, im=30, eps=1e-8, maxits=500):
    """kronecker.f90:181-365.  prec = (alu, jlu, ju) ILUT factors, or a BandFactor
    (exact K^-1).  Returns (sol, ierr, its, hist)."""
    q = A.shape[0]
    p = len(K[0]) - 1
    rhs = _f64(rhs).copy()
    sol = np.zeros(p * q) if sol0 is None else _f64(sol0).copy()
    cap = maxits + 64
    hist = np.zeros(cap)
    nh = C.c_int(0)
    its = C.c_int(0)
    args = (_fmat(A), _fmat(B), q, _i32(K[0]), _i32(K[1]), _f64(K[2]),
            _i32(M[0]), _i32(M[1]), _f64(M[2]), p)
    tail = (rhs, sol, im, eps, maxits, hist, cap, C.byref(nh), C.byref(its))
    if isinstance(prec, BandFactor):
        ierr = lib().or_kronpgmres_band(*args, prec.lb, prec.w, *tail)
    elif isinstance(prec, LapackBandFactor):
        ierr = lib().or_kronpgmres_cb(*args, prec.callback, None, *tail)
    else:
        alu, jlu, ju = prec
        ierr = lib().or_kronpgmres(*args, _f64(alu), _i32(jlu), _i32(ju), *tail)
    return sol, ierr, its.value, hist[:nh.value].copy()


class LapackBandFactor:
    """K^-1 for or_kronpgmres_cb: band Cholesky of an SPD CSR matrix by LAPACK (dpbtrf /
    dpbtrs through scipy), for meshes too large for or_bandchol's scalar loop.  The same
    exact factor (the live path's K^-1, up to rounding).  The solve is called back from
    the oracle's GMRES for the q temporal blocks at once."""

    def __init__(self, K):
        import scipy.linalg as sla
        ia, ja, a = _i32(K[0]), _i32(K[1]), _f64(K[2])
        n = len(ia) - 1
        rows = np.repeat(np.arange(n), np.diff(ia))
        cols = ja - 1
        low = cols <= rows
        self.n = n
        self.w = int(np.max(rows[low] - cols[low])) if n else 0
        ab = np.zeros((self.w + 1, n))
        ab[rows[low] - cols[low], cols[low]] = a[low]
        self.cb = sla.cholesky_banded(ab, lower=True, overwrite_ab=True, check_finite=False)
        self._sla = sla

        def solve(_ctx, p, q, x, w):
            X = np.ctypeslib.as_array(x, shape=(q, p))
            W = np.ctypeslib.as_array(w, shape=(q, p))
            W[:] = self._sla.cho_solve_banded((self.cb, True), X.T, check_finite=False).T
            return 0
        self.callback = SPATIAL_FN(solve)


class BandFactor:
    """Band Cholesky K = L L^T of an SPD CSR matrix (lower band, 1-based CSR)."""

    def __init__(self, K):
        ia, ja, a = _i32(K[0]), _i32(K[1]), _f64(K[2])
        n = len(ia) - 1
        rows = np.repeat(np.arange(n), np.diff(ia))
        cols = ja - 1
        low = cols <= rows
        self.n = n
        self.w = int(np.max(rows[low] - cols[low])) if n else 0
        ld = self.w + 1
        self.lb = np.zeros(ld * n)
        self.lb[(rows[low] - cols[low]) + cols[low] * ld] = a[low]
        self.info = lib().or_bandchol(n, self.w, self.lb)

    def solve(self, y):
        x = np.zeros(self.n)
        lib().or_bandsol(self.n, self.w, self.lb, _f64(y), x)
        return x

    def kronsol(self, A, x):
        """(A (x) K)^-1 x in kronlusol's order."""
        q = A.shape[0]
        y = np.zeros(self.n * q)
        if lib().or_kronbandsol(_fmat(A), q, self.lb, self.w, self.n, _f64(x), y):
            raise np.linalg.LinAlgError("Matrix is numerically singular!")
        return y


def stcoeff_m(t1, dt, w):
    """stcoeff.f90:62-356 -> dict of 6x6 arrays (k, m, ktn, mtn, ktn1, mtn1)."""
    outs = [np.zeros(36) for _ in range(6)]
    lib().or_stcoeff_m(t1, dt, w, *outs)
    names = ("k", "m", "ktn", "mtn", "ktn1", "mtn1")
    return {n: o.reshape(6, 6).T.copy() for n, o in zip(names, outs)}


def stcoeff_t(t1, dt):
    """stcoeff.f90:1-60 -> dict of 3x3 arrays."""
    outs = [np.zeros(9) for _ in range(6)]
    lib().or_stcoeff_t(t1, dt, *outs)
    names = ("k", "m", "ktn", "mtn", "ktn1", "mtn1")
    return {n: o.reshape(3, 3).T.copy() for n, o in zip(names, outs)}


def getfextc_m(t1, w, dt):
    """getfextc.f90:22-58 -> (fextc[6], fint[6])."""
    f = np.zeros(6)
    g = np.zeros(6)
    lib().or_getfextc_m(t1, w, dt, f, g)
    return f, g


def damage_nt(tcoorde, w, nip):
    """kokkos_damage.f90:36-47 -> Nt[nip, 6]."""
    out = np.zeros(6 * nip)
    lib().or_damage_nt(_f64(tcoorde), w, nip, out)
    return out.reshape(nip, 6)


def damage_params(params=None):
    """Set the damage-law constants (sig_f, sig_u, eps_D, Cy, h, S, s) of the oracle's
    functor; None restores the reference's (damage.h:9-17)."""
    if params is None:
        lib().or_damage_set_params(None)
    else:
        p = np.ascontiguousarray(params, dtype=np.float64)
        lib().or_damage_set_params(p.ctypes.data)


def kokkos_host(ne, nn, nip, nodes, Nx, Ny, Nz, Nt, sol, ym, nu, state):
    """kokkos_host.cpp:37-190 / kokkos_kernel.h:44-344 on the CPU.

    state: dict with eps_old, eps_pl, sig_eff, sig_back ([ngp*6]) and p, D, ws,
    del_po, sig_d ([ngp]); a new dict with updated copies is returned.
    """
    st = {k: _f64(v).copy() for k, v in state.items()}
    lib().or_kokkos_host(0, ne, nn, 3, 8, 3, 8, 6, nip, _i32(nodes), _f64(Nx), _f64(Ny),
                         _f64(Nz), _f64(Nt), _f64(sol), _f64(ym), _f64(nu),
                         st["eps_old"], st["eps_pl"], st["sig_eff"], st["sig_back"],
                         st["p"], st["D"], st["ws"], st["del_po"], st["sig_d"])
    return st


def damage_step(strain, ym, nu, eo, ep, se, sb, sc):
    """One return-map step (kokkos_kernel.h:85-341) on copies of the state; returns
    (eo, ep, se, sb, sc) with sc = (p, D, ws, del_po, sig_d)."""
    eo, ep, se, sb, sc = (_f64(v).copy() for v in (eo, ep, se, sb, sc))
    lib().or_damage_step(_f64(strain), float(ym), float(nu), eo, ep, se, sb, sc)
    return eo, ep, se, sb, sc


def chkelesta(ne, ngpe, Dc, D, elesta=None, gpsta=None):
    """delelement.f90:109-136 -> (nefail list 1-based, elesta, gpsta)."""
    el = np.ones(ne, dtype=np.int32) if elesta is None else _i32(elesta).copy()
    gp = np.ones(ne * ngpe, dtype=np.int32) if gpsta is None else _i32(gpsta).copy()
    nefail = np.zeros(ne, dtype=np.int32)
    nef = lib().or_chkelesta(ne, ngpe, Dc, _f64(D), el, gp, nefail)
    return nefail[:nef].copy(), el, gp


# ---- element deletion (delelement.f90, main3d.f90:650-686), numpy restatement --------

_HEX_SIGNS = np.array([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1],
                       [1, -1, 1], [1, -1, -1], [1, 1, -1], [1, 1, 1]], dtype=np.float64)


def hex_stiffness(X, C6):
    """Stiffness of one C3D8 hexahedron with 2x2x2 Gauss points (elekmtx,
    elesubs.f90:207-251): sum over points of B^T C6 B det J.  X [8, 3] node
    coordinates in the local order of shape3x (procs.f90:1320-1327); DOF 3a + c.
    Written from the mathematics (numpy products), so it agrees with the reference
    to rounding, not bit for bit."""
    g = 0.577350269189626
    ke = np.zeros((24, 24))
    for xi in (-g, g):
        for eta in (-g, g):
            for zeta in (-g, g):
                pt = np.array([xi, eta, zeta])
                f = 1.0 + _HEX_SIGNS * pt                      # [8, 3]
                dN = np.empty((3, 8))
                for m in range(3):
                    o = [k for k in range(3) if k != m]
                    dN[m] = 0.125 * _HEX_SIGNS[:, m] * f[:, o[0]] * f[:, o[1]]
                J = dN @ X
                G = np.linalg.solve(J, dN)                     # dN/dx [3, 8]
                B = np.zeros((6, 24))
                B[0, 0::3] = G[0]; B[1, 1::3] = G[1]; B[2, 2::3] = G[2]
                B[3, 0::3] = G[1]; B[3, 1::3] = G[0]
                B[4, 1::3] = G[2]; B[4, 2::3] = G[1]
                B[5, 0::3] = G[2]; B[5, 2::3] = G[0]
                ke += B.T @ C6 @ B * np.linalg.det(J)
    return ke


def deletion_block(xconn, xcoord, K, M, fixed, elesta, nefail, C6):
    """main3d.f90:650-686 restated on dense-free CSR: K minus the failed elements'
    stiffness, survivor counts by brute force over the still-active elements,
    orphan nodes (those of the failed elements with no active element left, in the
    order the failed list reaches them), unit diagonal and fixed DOFs on them, the
    1-0 method on K and M.  elesta: 1 for elements active before this event.
    Returns (K', M', fixed', ndfail) with K', M' as scipy CSR."""
    import scipy.sparse as sp
    n = len(K[0]) - 1
    csr = lambda T: sp.csr_matrix((np.asarray(T[2]), np.asarray(T[1]) - 1, np.asarray(T[0]) - 1),
                                  shape=(n, n))
    Kd = csr(K).tolil()
    for e in nefail:
        nodes = xconn[e - 1] - 1
        dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        Kd[np.ix_(dofs, dofs)] = Kd[np.ix_(dofs, dofs)].toarray() - hex_stiffness(xcoord[nodes], C6)
    alive = np.asarray(elesta, dtype=bool).copy()
    alive[np.asarray(nefail) - 1] = False
    count = np.zeros(len(xcoord), dtype=np.int64)
    for e in np.flatnonzero(alive):
        count[xconn[e] - 1] += 1
    ndfail = []
    for e in nefail:
        for node in xconn[e - 1]:
            if count[node - 1] == 0 and node not in ndfail:
                ndfail.append(int(node))
    fixed = np.array(fixed, dtype=np.int32)
    for node in ndfail:
        for c in range(3):
            Kd[3 * (node - 1) + c, 3 * (node - 1) + c] += 1.0
        fixed[3 * (node - 1):3 * node] = 1

    def bc(S):
        S = sp.csr_matrix(S)
        keep = sp.diags((fixed == 0).astype(np.float64))
        out = keep @ S @ keep + sp.diags((fixed != 0).astype(np.float64))
        out.eliminate_zeros()
        return out.tocsr()

    return bc(Kd.tocsr()), bc(csr(M)), fixed, np.array(ndfail, dtype=np.int32)


def thermo_slabs(mesh, fixed_m, tfixed, rext_m, rext_t, nslab, dt, w, ym0, nu0, rho, kappa,
                 cp, alpha, Tref, T0, mechanical=True):
    """The thermo-mechanical slab loop of main3d.f90:579-834 (damage and deletion left
    out) with exact sparse direct solves in place of the preconditioned GMRES.

    Setup: K_t, M_t with and without the 1-0 method (main3d.f90:365-453), fInit =
    (A_t (x) K_t0 + B_t (x) M_t0) T_ref and fBound = -(...) T_fixed (main3d.f90:603-608).
    Per slab: thermal right-hand side (fInit on the first slab, fextc_t(j) rext_t, the
    jump (ktn1 (x) K_t + mtn1 (x) M_t) T_prev by kronspgemv, fBound, zero at fixed
    DOFs), solve, T0 at fixed DOFs (main3d.f90:699-731); material_update and the
    thermal load vv from the third temporal node, re-assembly with per-Gauss-point E and
    nu, 1-0 method (main3d.f90:752-781); mechanical right-hand side fextc_m(j) rext_m +
    fint_m(j) vv plus the jump, zero at fixed DOFs, solve, zero at fixed DOFs
    (main3d.f90:803-840).  The slab matrices use t1 = 0 (main3d.f90:717, 817 pass the
    integer slab index to the real t1, a subnormal; stcoeff_t ignores it).

    The spatial matrices, the material update and vv come from the host restatements
    in xtfem_amd.fem, which the CPU tests pin bit for bit against the compiled
    reference; the temporal data, the Kronecker products and the solves are this
    module's.  Returns a list of (tempst [3*nn], dispst [18*nn], ym, nu, vv) per slab,
    temporal-major (only tempst with ``mechanical=False``)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from xtfem_amd import fem, slab
    nn = mesh.nn
    p = 3 * nn
    csr = lambda S, n: sp.csr_matrix((S[2], S[1] - 1, S[0] - 1), shape=(n, n))
    Kt0, Mt0 = fem.assemble_t(mesh, kappa, cp, rho)
    tf = np.asarray(tfixed) != 0
    Kt, Mt = fem.applybc(Kt0, tf.astype(np.int32)), fem.applybc(Mt0, tf.astype(np.int32))
    fx_t = np.tile(tf, 3)
    fx0_t = np.where(fx_t, T0, 0.0)
    fx_m = np.tile(np.asarray(fixed_m) != 0, 6)
    ct = stcoeff_t(0.0, dt)
    At, Bt = ct["k"] + ct["ktn"], ct["m"] + ct["mtn"]
    tempst = np.full(3 * nn, float(Tref))
    finit = kronspgemv(At, Bt, 1.0, 1.0, Kt0, Mt0, tempst, np.zeros(3 * nn))
    tempst = fx0_t.copy()
    fbound = kronspgemv(-At, -Bt, 1.0, 1.0, Kt0, Mt0, tempst, np.zeros(3 * nn))
    lu_t = spl.splu((sp.kron(At, csr(Kt, nn)) + sp.kron(Bt, csr(Mt, nn))).tocsc())
    cm = stcoeff_m(0.0, dt, w)
    Am, Bm = cm["k"] + cm["ktn"], cm["m"] + cm["mtn"]
    dispst = np.zeros(6 * p)
    tc = np.zeros(3)
    out = []
    for ii in range(1, nslab + 1):
        tc = tc[2] + np.arange(3) * dt / 2
        f = finit.copy() if ii == 1 else np.zeros(3 * nn)
        fe = slab.getfextc_t(tc[0], dt)
        for j in range(3):
            f[j * nn:(j + 1) * nn] += fe[j] * rext_t
        f = kronspgemv(ct["ktn1"], ct["mtn1"], 1.0, 1.0, Kt, Mt, tempst, f)
        f = f + fbound
        f[fx_t] = 0.0
        tempst = lu_t.solve(f)
        tempst[fx_t] = fx0_t[fx_t]
        T = tempst[2 * nn:]
        if not mechanical:
            out.append((tempst.copy(), None, None, None, None))
            continue
        ym, nu = fem.material_update(mesh, T, Tref, ym0, nu0)
        vv = fem.thermal_load(mesh, T, Tref, alpha, ym, nu)
        K, M, *_ = fem.assemble(mesh, ym, nu, rho)
        Kb, Mb = fem.applybc(K, fixed_m), fem.applybc(M, fixed_m)
        fe_m, fi_m = getfextc_m(tc[0], w, dt)
        g = np.concatenate([fe_m[j] * rext_m + fi_m[j] * vv for j in range(6)])
        g = kronspgemv(cm["ktn1"], cm["mtn1"], 1.0, 1.0, Kb, Mb, dispst, g)
        g[fx_m] = 0.0
        S = (sp.kron(Am, csr(Kb, p)) + sp.kron(Bm, csr(Mb, p))).tocsc()
        dispst = spl.spsolve(S, g)
        dispst[fx_m] = 0.0
        out.append((tempst.copy(), dispst.copy(), ym, nu, vv))
    return out
