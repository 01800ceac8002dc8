"""ctypes front end of the reference build ``oracle/_ref/libxtfem_ref.so``.

TEST INFRASTRUCTURE ONLY.  The library is the reference's own Fortran (kronecker.f90,
ilut.f, stcoeff.f90, getfextc.f90, getdfinal.f and the FEM setup routines of
elesubs.f90 / spmtx.f90 / mpisubs.f90 / procs.f90) compiled in place with AMD flang
by ``make -C oracle ref``.  It exists only where /root/reference is mounted; every
caller must treat it as optional (``available()``).

All Fortran arguments are passed by reference; arrays are Fortran-ordered.
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libxtfem_ref.so")
_LIB = None


def available() -> bool:
    return os.path.exists(_PATH)


def lib():
    global _LIB
    if _LIB is None:
        os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
        _LIB = C.CDLL(_PATH)
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ri(v):
    return C.byref(C.c_int(int(v)))


def _rd(v):
    return C.byref(C.c_double(float(v)))


def _F(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def stcoeff_m(t1, dt, w):
    outs = [np.zeros((6, 6), order="F") for _ in range(6)]
    lib().stcoeff_m_(_rd(t1), _rd(dt), _rd(w), *[_p(o) for o in outs])
    return dict(zip(("k", "m", "ktn", "mtn", "ktn1", "mtn1"), [np.array(o) for o in outs]))


def stcoeff_t(t1, dt):
    outs = [np.zeros((3, 3), order="F") for _ in range(6)]
    lib().stcoeff_t_(_rd(t1), _rd(dt), *[_p(o) for o in outs])
    return dict(zip(("k", "m", "ktn", "mtn", "ktn1", "mtn1"), [np.array(o) for o in outs]))


def getfextc_m(t1, w, dt):
    f = np.zeros(6)
    g = np.zeros(6)
    lib().getfextc_m_(_rd(t1), _rd(w), _rd(dt), _p(f), _p(g))
    return f, g


def amux(ia, ja, a, x):
    ia = _F(ia, np.int32); ja = _F(ja, np.int32); a = _F(a); x = _F(x)
    n = len(ia) - 1
    y = np.zeros(n)
    lib().amux_(_ri(n), _p(x), _p(y), _p(a), _p(ja), _p(ia))
    return y


def _csr(S):
    return _F(S[0], np.int32), _F(S[1], np.int32), _F(S[2])


def kronspmv(A, B, K, M, x):
    q = A.shape[0]
    ki, kj, kv = _csr(K)
    mi, mj, mv = _csr(M)
    p = len(ki) - 1
    x = _F(x).copy()
    y = np.zeros(p * q)
    lib().kronspmv_(_p(_F(A)), _p(_F(B)), _ri(q), _p(ki), _p(kj), _p(kv), _p(mi), _p(mj),
                    _p(mv), _ri(p), _p(x), _p(y))
    return y


def kronspgemv(A, B, alpha, beta, K, M, x, y):
    q = A.shape[0]
    ki, kj, kv = _csr(K)
    mi, mj, mv = _csr(M)
    p = len(ki) - 1
    x = _F(x).copy()
    y = _F(y).copy()
    lib().kronspgemv_(_p(_F(A)), _p(_F(B)), _rd(alpha), _rd(beta), _ri(q), _p(ki), _p(kj),
                      _p(kv), _p(mi), _p(mj), _p(mv), _ri(p), _p(x), _p(y))
    return y


def ilut(K, lfil, droptol, iwk=None):
    ia, ja, a = _csr(K)
    n = len(ia) - 1
    if iwk is None:
        iwk = n + 2 + n * (2 * lfil + 1) + 16
    alu = np.zeros(iwk)
    jlu = np.zeros(iwk, dtype=np.int32)
    ju = np.zeros(n, dtype=np.int32)
    w = np.zeros(n + 1)
    jw = np.zeros(2 * n, dtype=np.int32)
    ierr = C.c_int(0)
    lib().ilut_(_ri(n), _p(a), _p(ja), _p(ia), _ri(lfil), _rd(droptol), _p(alu), _p(jlu),
                _p(ju), _ri(iwk), _p(w), _p(jw), C.byref(ierr))
    nnz = int(jlu[n]) - 1 if ierr.value == 0 else 0
    return alu[:max(nnz, n + 1)].copy(), jlu[:max(nnz, n + 1)].copy(), ju, ierr.value


def lusol(alu, jlu, ju, y):
    alu = _F(alu); jlu = _F(jlu, np.int32); ju = _F(ju, np.int32); y = _F(y).copy()
    n = len(ju)
    x = np.zeros(n)
    lib().lusol_(_ri(n), _p(y), _p(x), _p(alu), _p(jlu), _p(ju))
    return x


def kronlusol(A, alu, jlu, ju, x):
    q = A.shape[0]
    alu = _F(alu); jlu = _F(jlu, np.int32); ju = _F(ju, np.int32)
    p = len(ju)
    x = _F(x).copy()
    y = np.zeros(p * q)
    lib().kronlusol_(_p(_F(A)), _ri(q), _p(alu), _p(jlu), _p(ju), _ri(p), _p(x), _p(y))
    return y


def kronpgmres(A, B, K, M, prec, rhs, sol0=None, im=30, eps=1e-8, maxits=500, log=False):
    """kronecker.f90:181-365.  With ``log`` the residual history the reference writes to
    its log unit (gmres.log, FORMAT 199) is returned too, as text."""
    if log:
        import tempfile
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as d:
            os.chdir(d)
            try:
                sol, ierr = _kronpgmres(A, B, K, M, prec, rhs, sol0, im, eps, maxits, 40)
                lib().xt_ref_close_(_ri(40))
            finally:
                os.chdir(cwd)
            name = os.path.join(d, "fort.40")
            text = open(name).read() if os.path.exists(name) else ""
        return sol, ierr, text
    return _kronpgmres(A, B, K, M, prec, rhs, sol0, im, eps, maxits, -1)


def _kronpgmres(A, B, K, M, prec, rhs, sol0, im, eps, maxits, iout):
    q = A.shape[0]
    ki, kj, kv = _csr(K)
    mi, mj, mv = _csr(M)
    p = len(ki) - 1
    alu = _F(prec[0]); jlu = _F(prec[1], np.int32); ju = _F(prec[2], np.int32)
    rhs = _F(rhs).copy()
    sol = np.zeros(p * q) if sol0 is None else _F(sol0).copy()
    ierr = C.c_int(0)
    lib().kronpgmres_(_p(_F(A)), _p(_F(B)), _ri(q), _p(ki), _p(kj), _p(kv), _p(mi), _p(mj),
                      _p(mv), _ri(p), _p(alu), _p(jlu), _p(ju), _p(rhs), _p(sol), _ri(im),
                      _rd(eps), _ri(maxits), _ri(iout), C.byref(ierr))
    return sol, ierr.value


def getdfinal(strain_tgo, strain_tg, strain_pl, p, D, ws, sig_eff, backsig, del_po, sig_d,
              cmat, dmat):
    """getdfinal.f:10-397, one return-map step.  Array arguments are updated in place;
    the scalars come back in a tuple (p, D, ws, del_po, sig_d)."""
    sc = [C.c_double(float(v)) for v in (p, D, ws, del_po, sig_d)]
    lib().getdfinal_(_p(strain_tgo), _p(strain_tg), _p(strain_pl), C.byref(sc[0]),
                     C.byref(sc[1]), C.byref(sc[2]), _p(sig_eff), _p(backsig), C.byref(sc[3]),
                     C.byref(sc[4]), _p(cmat), _p(dmat))
    return tuple(s.value for s in sc)


def chkelesta(nn, ne, ngpe, Dc, D):
    elesta = np.ones(ne, dtype=np.int32)
    gpsta = np.ones(ne * ngpe, dtype=np.int32)
    nefail = np.zeros(ne, dtype=np.int32)
    nef = C.c_int(0)
    D = _F(D)
    lib().chkelesta_(_ri(nn), _ri(ne), _ri(8), _ri(ngpe), _rd(Dc), _p(D), _p(elesta),
                     _p(gpsta), C.byref(nef), _p(nefail))
    return nefail[:nef.value].copy(), elesta, gpsta


def assemble_m(xcoord, xconn, ym0, nu0, rho, ym=None, nu=None, temp=None, alpha=0.0,
               Tref=0.0, with_vv=False):
    """mpiglbmtx_m (mpisubs.f90:69-139) over all elements, then tricsr
    (spmtx.f90:359-407).  Returns 1-based CSR triples K, M and Nx, Ny, Nz [8, 8*ne];
    with ``with_vv`` also the thermal load vv [3*nn] that the same call accumulates
    from the nodal temperatures ``temp`` (calc_thermo_stress + elevec_m).  ym, nu:
    per-Gauss-point arrays [8*ne] (default ym0, nu0 everywhere)."""
    L = lib()
    xcoord = _F(xcoord)
    xconn = _F(xconn, np.int32)
    nn = xcoord.shape[0]
    ne = xconn.shape[0]
    ndofn, nnse, nquads, ngpe, ncmp, nnsd = 3, 8, 2, 8, 6, 2
    wqs = np.zeros(nquads)
    ns = np.zeros((ndofn * ngpe, ndofn * nnse), order="F")
    dns = np.zeros((ndofn * ngpe, nnse), order="F")
    dnxyz = np.zeros((ndofn ** 2 * ngpe, ndofn * nnse), order="F")
    L.setshapefunc_m_(_ri(ndofn), _ri(nnsd), _ri(nnse), _ri(nquads), _ri(ngpe), _p(wqs),
                      _p(ns), _p(dns), _p(dnxyz))
    wqs_t = np.zeros(nquads)
    ns_t = np.zeros((ngpe, nnse), order="F")
    dns_t = np.zeros((ngpe, nnse, 3), order="F")
    L.setshapefunc_t_(_ri(nnsd), _ri(nnse), _ri(nquads), _ri(ngpe), _p(wqs_t), _p(ns_t),
                      _p(dns_t))
    ele = np.arange(1, ne + 1, dtype=np.int32)
    ym = np.full(ngpe * ne, float(ym0)) if ym is None else _F(ym)
    nu = np.full(ngpe * ne, float(nu0)) if nu is None else _F(nu)
    sd = np.zeros((6, 9), order="F")
    stiff = np.zeros((6, 6, ne * ngpe), order="F")
    Nx = np.zeros((nnse, ngpe * ne), order="F")
    Ny = np.zeros_like(Nx)
    Nz = np.zeros_like(Nx)
    knz = ne * (nnse * ndofn) ** 2
    ndof = nn * ndofn
    nti = max(knz, ndof + 1)
    ki = np.zeros(nti, dtype=np.int32); kj = np.zeros(knz, dtype=np.int32); kv = np.zeros(knz)
    mi = np.zeros(nti, dtype=np.int32); mj = np.zeros(knz, dtype=np.int32); mv = np.zeros(knz)
    vv = np.zeros(ndof)
    tempst = np.zeros(nn) if temp is None else _F(temp)
    L.mpiglbmtx_m_(_ri(ne), _p(ele), _ri(ne), _ri(nn), _ri(ndofn), _ri(nnse), _ri(nquads),
                   _ri(ngpe), _ri(ncmp), _p(xconn), _p(xcoord), _p(ym), _p(nu), _rd(ym0),
                   _rd(nu0), _rd(rho), _rd(alpha), _rd(Tref), _p(wqs), _p(ns_t), _p(ns), _p(dns_t),
                   _p(dns), _p(dnxyz), _p(sd), _p(tempst), _p(stiff), _p(Nx), _p(Ny), _p(Nz),
                   _p(ki), _p(kj), _p(kv), _ri(knz), _p(mi), _p(mj), _p(mv), _ri(knz), _p(vv))
    out = []
    for (ti, tj, tv) in ((ki, kj, kv), (mi, mj, mv)):
        nz = C.c_int(knz)
        L.tricsr_(_p(ti), _p(tj), _p(tv), _ri(ndof), C.byref(nz), _ri(nti))
        out.append((ti[:ndof + 1].copy(), tj[:nz.value].copy(), tv[:nz.value].copy()))
    if with_vv:
        return out[0], out[1], np.array(Nx), np.array(Ny), np.array(Nz), vv.copy()
    return out[0], out[1], np.array(Nx), np.array(Ny), np.array(Nz)


def shape_tables_t(nquads=2):
    """setshapefunc_t (elesubs.f90:337-371): wqs [nquads], ns [ngpe, 8], dns [ngpe, 8, 3]."""
    ngpe = nquads ** 3
    wqs = np.zeros(nquads)
    ns = np.zeros((ngpe, 8), order="F")
    dns = np.zeros((ngpe, 8, 3), order="F")
    lib().setshapefunc_t_(_ri(2), _ri(8), _ri(nquads), _ri(ngpe), _p(wqs), _p(ns), _p(dns))
    return wqs, np.array(ns), np.array(dns)


def assemble_t(xcoord, xconn, kappa, cp, rho):
    """mpiglbmtx_t (mpisubs.f90:141-188, elemtx_t elesubs.f90:60-116) over all
    elements, then tricsr: thermal conductivity K_t and capacity M_t (1 DOF a node)."""
    L = lib()
    xcoord = _F(xcoord)
    xconn = _F(xconn, np.int32)
    nn, ne = xcoord.shape[0], xconn.shape[0]
    wqs, ns, dns = (_F(a) for a in shape_tables_t())
    ele = np.arange(1, ne + 1, dtype=np.int32)
    knz = ne * 64
    nti = max(knz, nn + 1)
    ki = np.zeros(nti, dtype=np.int32); kj = np.zeros(knz, dtype=np.int32); kv = np.zeros(knz)
    mi = np.zeros(nti, dtype=np.int32); mj = np.zeros(knz, dtype=np.int32); mv = np.zeros(knz)
    L.mpiglbmtx_t_(_ri(ne), _p(ele), _ri(ne), _ri(nn), _ri(1), _ri(8), _ri(2), _ri(8),
                   _p(xconn), _p(xcoord), _rd(kappa), _rd(cp), _rd(rho), _p(wqs), _p(ns),
                   _p(dns), _p(ki), _p(kj), _p(kv), _ri(knz), _p(mi), _p(mj), _p(mv), _ri(knz))
    out = []
    for (ti, tj, tv) in ((ki, kj, kv), (mi, mj, mv)):
        nz = C.c_int(knz)
        L.tricsr_(_p(ti), _p(tj), _p(tv), _ri(nn), C.byref(nz), _ri(nti))
        out.append((ti[:nn + 1].copy(), tj[:nz.value].copy(), tv[:nz.value].copy()))
    return out[0], out[1]


def material_update(xconn, nn, tempst, Tref, ym0, nu0):
    """thermoelastic.f90:1-46: E and nu at every Gauss point from the temperature of
    the slab's third temporal node, tempst [3*nn] (temporal-major)."""
    xconn = _F(xconn, np.int32)
    ne = xconn.shape[0]
    _, ns, _ = shape_tables_t()
    ele = np.arange(1, ne + 1, dtype=np.int32)
    tempst = _F(tempst)
    ym = np.zeros(8 * ne)
    nu = np.zeros(8 * ne)
    lib().material_update_(_ri(8), _ri(2), _ri(len(tempst)), _ri(nn), _ri(ne), _ri(8),
                           _ri(ne), _p(ele), _p(xconn), _p(_F(ns)), _p(tempst), _rd(Tref),
                           _rd(ym0), _rd(nu0), _p(ym), _p(nu))
    return ym, nu


def getfextc_t(t1, dt):
    """getfextc.f90:1-20 -> fextc [3]."""
    f = np.zeros(3)
    lib().getfextc_t_(_rd(t1), _rd(dt), _p(f))
    return f


def csrapplybc(S, bc):
    """spmtx.f90:268-307 (1-0 method, zero removal)."""
    ia, ja, a = (np.array(S[0], dtype=np.int32), np.array(S[1], dtype=np.int32),
                 np.array(S[2], dtype=np.float64))
    n = len(ia) - 1
    nz = C.c_int(len(a))
    bc = np.asarray(bc, dtype=np.int32)
    lib().csrapplybc_(_p(ia), _p(ja), _p(a), _ri(n), C.byref(nz), _p(bc))
    return ia, ja[:nz.value].copy(), a[:nz.value].copy()


def shape_tables_m(nquads=2):
    """setshapefunc_m (elesubs.f90:299-335): wqs [nquads], dns [3*ngpe, 8], dnxyz [9*ngpe, 24]."""
    ngpe = nquads ** 3
    wqs = np.zeros(nquads)
    ns = np.zeros((3 * ngpe, 24), order="F")
    dns = np.zeros((3 * ngpe, 8), order="F")
    dnxyz = np.zeros((9 * ngpe, 24), order="F")
    lib().setshapefunc_m_(_ri(3), _ri(2), _ri(8), _ri(nquads), _ri(ngpe), _p(wqs), _p(ns),
                          _p(dns), _p(dnxyz))
    return wqs, np.array(dns), np.array(dnxyz)


def setsdstiff(ym, nu):
    """elesubs.f90:374-401 -> (sd [6, 9], stiff [6, 6])."""
    sd = np.zeros((6, 9), order="F")
    stiff = np.zeros((6, 6), order="F")
    lib().setsdstiff_(_rd(ym), _rd(nu), _p(sd), _p(stiff))
    return np.array(sd), np.array(stiff)


def _csr_room(K, extra):
    ia = _F(K[0], np.int32).copy()
    nz = int(ia[-1]) - 1
    ja = np.zeros(nz + extra, dtype=np.int32)
    a = np.zeros(nz + extra)
    ja[:nz] = K[1][:nz]
    a[:nz] = K[2][:nz]
    return ia, ja, a, C.c_int(nz)


def delelement(xconn, xcoord, dns, dnxyz, sd, stiff, wqs, nefail, K):
    """delelement.f90:1-59 -> new K (1-based CSR)."""
    xconn = _F(xconn, np.int32)
    xcoord = _F(xcoord)
    ne, nnse = xconn.shape
    nn = xcoord.shape[0]
    nquads = len(wqs)
    nefail = _F(nefail, np.int32)
    ki, kj, kv, knz = _csr_room(K, len(nefail) * (3 * nnse) ** 2)
    ndof = len(ki) - 1
    lib().delelement_(_ri(nn), _ri(ne), _ri(3), _ri(nnse), _ri(nquads ** 3), _ri(6),
                      _ri(nquads), _p(xconn), _p(xcoord), _p(_F(dns)), _p(_F(dnxyz)),
                      _p(_F(sd)), _p(_F(stiff)), _p(_F(wqs)), _ri(len(nefail)), _p(nefail),
                      _p(ki), _p(kj), _p(kv), _ri(ndof), C.byref(knz))
    return ki, kj[:knz.value].copy(), kv[:knz.value].copy()


def chkcnsta(nn, xconn, nefail, necnt):
    """delelement.f90:139-161 -> updated necnt."""
    xconn = _F(xconn, np.int32)
    necnt = _F(necnt, np.int32).copy()
    nefail = _F(nefail, np.int32)
    lib().chkcnsta_(_ri(nn), _ri(xconn.shape[0]), _ri(xconn.shape[1]), _p(xconn),
                    _ri(len(nefail)), _p(nefail), _p(necnt))
    return necnt


def chkndsta(nn, xconn, nefail, necnt):
    """delelement.f90:164-201 -> orphan node ids (1-based, discovery order)."""
    xconn = _F(xconn, np.int32)
    nefail = _F(nefail, np.int32)
    necnt = _F(necnt, np.int32)
    ndfail = np.zeros(nn, dtype=np.int32)
    ndf = C.c_int(0)
    lib().chkndsta_(_ri(nn), _ri(xconn.shape[0]), _ri(xconn.shape[1]), _p(xconn),
                    _ri(len(nefail)), _p(nefail), _p(necnt), C.byref(ndf), _p(ndfail))
    return ndfail[:ndf.value].copy()


def adddiagonal(ndofn, ndfail, K):
    """delelement.f90:204-247 -> K + I on the orphan DOFs.  (Its tricsr call passes
    five of tricsr's six arguments; the sixth only sizes a dummy array.)"""
    ndfail = _F(ndfail, np.int32)
    ki, kj, kv, knz = _csr_room(K, len(ndfail) * ndofn)
    ndof = len(ki) - 1
    lib().adddiagonal_(_ri(8), _ri(ndofn), _ri(len(ndfail)), _p(ndfail), _p(ki), _p(kj), _p(kv),
                       _ri(ndof), C.byref(knz))
    return ki, kj[:knz.value].copy(), kv[:knz.value].copy()


def deletion_block(xconn, xcoord, K, M, fixed, necnt, nefail, ym, nu):
    """The reference's deletion block main3d.f90:650-686 on one rank, run through the
    compiled routines: delelement, chkcnsta, chkndsta, adddiagonal, the orphan
    constraints, csrapplybc on M and K.  K, M: current (constrained) matrices.
    Returns (K', M', fixed', necnt', ndfail)."""
    wqs, dns, dnxyz = shape_tables_m()
    sd, stiff = setsdstiff(ym, nu)
    K1 = delelement(xconn, xcoord, dns, dnxyz, sd, stiff, wqs, nefail, K)
    necnt = chkcnsta(len(xcoord), xconn, nefail, necnt)
    ndfail = chkndsta(len(xcoord), xconn, nefail, necnt)
    K2 = adddiagonal(3, ndfail, K1)
    fixed = np.array(fixed, dtype=np.int32)
    for j in ndfail:
        fixed[3 * (j - 1):3 * j] = 1
    return csrapplybc(K2, fixed), csrapplybc(M, fixed), fixed, necnt, ndfail


def dispinterp(nodes, sol, Nt, nn):
    """postprocess.f90:54-76 -> disp [8*3] of one element at one temporal point:
    nodes (1-based, 8), sol temporal-major [6*3*nn], Nt [6]."""
    disp = np.zeros(24)
    lib().dispinterp_(_ri(3), _ri(8), _ri(nn), _ri(3), _p(_F(nodes, np.int32)), _p(_F(sol)),
                      _p(_F(Nt)), _p(disp))
    return disp


def calstrain(disp, Nx, Ny, Nz, dmat):
    """postprocess.f90:79-104 -> (strain [6], stress [6]) at one Gauss point."""
    strain = np.zeros(6)
    stress = np.zeros(6)
    lib().calstrain_(_ri(8), _ri(3), _ri(6), _p(_F(disp)), _p(_F(Nx)), _p(_F(Ny)), _p(_F(Nz)),
                     _p(_F(dmat)), _p(strain), _p(stress))
    return strain, stress


def mpipostprocess(job, xconn, Nx, Ny, Nz, ym, nu, w, tcoorde, sol, tempst, gpsta, elsta, D,
                   ws, nip, nipc):
    """postprocess.f90:106-210 run in a scratch directory (rank 0): returns the text of
    every file it appends to, keyed by suffix (time, disp, temp, strain, stress, damage,
    energy, status).  Nx/Ny/Nz: [8, 8*ne] (column = Gauss point)."""
    import tempfile
    xconn = _F(xconn, np.int32)
    ne = xconn.shape[0]
    nn = len(tempst) // 3
    name = job.ljust(80).encode()
    cwd = os.getcwd()
    out = {}
    with tempfile.TemporaryDirectory() as d:
        os.chdir(d)
        try:
            lib().mpipostprocess_(C.c_char_p(name), _ri(ne), _ri(nn), _ri(3), _ri(8), _ri(3),
                                  _ri(8), _ri(6), _ri(nip), _ri(nipc), _p(xconn), _p(_F(Nx)),
                                  _p(_F(Ny)), _p(_F(Nz)), _p(_F(ym)), _p(_F(nu)), _rd(w),
                                  _p(_F(tcoorde)), _p(_F(sol)), _p(_F(tempst)),
                                  _p(_F(gpsta, np.int32)), _p(_F(elsta, np.int32)), _p(_F(D)),
                                  _p(_F(ws)), _ri(0), C.c_size_t(80))
        finally:
            os.chdir(cwd)
        for f in os.listdir(d):
            out[f.rsplit(".", 1)[1]] = open(os.path.join(d, f)).read()
    return out
