"""Result files of the driver, in the reference's formats (postprocess.f90:1-469) and the
GMRES residual log (kronecker.f90:257-258, 311, 354).

Output step: every ``nout`` slabs (main3d.f90:913-917) the driver interpolates the
slab's solution at one temporal point, index nout = nip - (nipc/4)*3 + 1 of the nip
points t_i = t1 + (i-1) dt / nip (postprocess.f90:140-152), and writes

* ``job.time``   the time of that point (rank 0),
* ``job.disp``   nodal displacements, one node a record (dispinterpt, rank 0),
* ``job.temp``   nodal temperatures, from the thermal slab's 3 temporal nodes
  (tempinterpt, rank 0),
* ``job_<rank>.strain`` / ``.stress``  6 components per Gauss point, '(6(E16.8,2X))',
  from dispinterp + calstrain with the elastic matrix of that point's E and nu (the
  shear terms as coef (1 - 2 nu) / 2); zero at failed Gauss points,
* ``job_<rank>.damage`` / ``.energy``  D and ws per Gauss point, ``.status`` the element
  status, list-directed.

All files are appended to (initfile only creates them).  Edit-descriptor records
(E16.8, D20.6, the log's I4) are reproduced character for character.  List-directed
records are compiler-defined (the reference builds with whatever mpif90 wraps); they
are written in the layout of the LLVM Fortran runtime (one blank before each item; F
form without a leading zero when |x| rounded to one digit lies in [0.1, 1e15), else one
digit before the point and a two- or three-digit exponent) with the shortest digits
that round-trip, so every item reads back as the same double.  The arithmetic follows the reference's loop
order: the numbers are bit-identical for the same inputs.
"""
from __future__ import annotations

import os

import numpy as np

from .fem import Mesh


def output_index(nip, nipc):
    """1-based temporal point written at an output step (postprocess.f90:141)."""
    return nip - nipc // 4 * 3 + 1


def temporal_functions(tcoorde, w, nip, i):
    """Nt(1:6) at point i (1-based) of the slab (postprocess.f90:142-150) and its time."""
    tc = np.asarray(tcoorde, dtype=np.float64)
    dt = tc[2] - tc[0]
    t = tc[0] + (i - 1) * dt / nip
    n1 = 2.0 * (tc[1] - t) * (tc[2] - t) / (dt * dt)
    n2 = -4.0 * (tc[0] - t) * (tc[2] - t) / (dt * dt)
    n3 = 2.0 * (tc[0] - t) * (tc[1] - t) / (dt * dt)
    s = np.sin(w * t)
    nt = np.array([n1, n2, n3, n1 * (s - np.sin(w * tc[0])), n2 * (s - np.sin(w * tc[1])),
                   n3 * (s - np.sin(w * tc[2]))])
    return nt, float(t)


def nodal_displacement(sol, nt, nn):
    """dispinterpt (postprocess.f90:32-52): disp [nn, 3] = sum_i Nt(i) u_i, temporal
    functions in order; sol temporal-major [6 * 3 * nn]."""
    u = np.asarray(sol, dtype=np.float64).reshape(6, nn, 3)
    d = np.zeros((nn, 3))
    for i in range(6):
        d = d + nt[i] * u[i]
    return d


def nodal_temperature(tempst, nt, nn):
    """tempinterpt (postprocess.f90:11-29): the 3 temporal nodes of the thermal slab
    weighted by Nt(1:3)."""
    T = np.asarray(tempst, dtype=np.float64).reshape(3, nn)
    t = np.zeros(nn)
    for i in range(3):
        t = t + nt[i] * T[i]
    return t


def gauss_strain_stress(mesh: Mesh, Nx, Ny, Nz, ym, nu, sol, nt, gpsta=None):
    """Strain and stress [8*ne, 6] at every Gauss point (postprocess.f90:153-188):
    dispinterp of the element's nodal values, calstrain (sums over the 8 nodes in order,
    engineering shear), stress = Dmat strain with Dmat(i,i) = coef (1 - nu),
    Dmat(i+3,i+3) = coef (1 - 2 nu) / 2, Dmat(i,j) = coef nu, coef = E / ((1 + nu)
    (1 - 2 nu)).  Nx/Ny/Nz: [8*ne, 8] (row = Gauss point).  Failed points (gpsta 0) are
    zero."""
    ne, nn = mesh.ne, mesh.nn
    u = np.asarray(sol, dtype=np.float64).reshape(6, nn, 3)
    nodes = mesh.xconn - 1                                        # [ne, 8]
    de = np.zeros((ne, 8, 3))
    for i in range(6):
        de = de + nt[i] * u[i][nodes]
    de = np.repeat(de, 8, axis=0)                                 # per Gauss point
    Nx, Ny, Nz = (np.asarray(a, dtype=np.float64).reshape(8 * ne, 8) for a in (Nx, Ny, Nz))
    e = np.zeros((8 * ne, 6))
    for a in range(8):
        d1, d2, d3 = de[:, a, 0], de[:, a, 1], de[:, a, 2]
        e[:, 0] = e[:, 0] + Nx[:, a] * d1
        e[:, 1] = e[:, 1] + Ny[:, a] * d2
        e[:, 2] = e[:, 2] + Nz[:, a] * d3
        e[:, 3] = e[:, 3] + Ny[:, a] * d1 + Nx[:, a] * d2
        e[:, 4] = e[:, 4] + Nz[:, a] * d2 + Ny[:, a] * d3
        e[:, 5] = e[:, 5] + Nz[:, a] * d1 + Nx[:, a] * d3
    E = np.broadcast_to(np.asarray(ym, dtype=np.float64), (8 * ne,))
    v = np.broadcast_to(np.asarray(nu, dtype=np.float64), (8 * ne,))
    coef = E / ((1.0 + v) * (1.0 - 2.0 * v))
    dd, od, sh = coef * (1.0 - v), coef * v, coef * (1.0 - 2.0 * v) / 2.0
    s = np.zeros_like(e)
    for i in range(3):
        acc = np.zeros(8 * ne)
        for j in range(3):
            acc = acc + (dd if i == j else od) * e[:, j]
        s[:, i] = acc
        s[:, i + 3] = sh * e[:, i + 3]
    if gpsta is not None:
        bad = np.asarray(gpsta) == 0
        e[bad] = 0.0
        s[bad] = 0.0
    return e, s


# ---------------------------------------------------------------------------------
# Formatting
# ---------------------------------------------------------------------------------

def _digits(x):
    """Shortest round-trip decimal digits (nearest of that length) and exponent:
    |x| = 0.DIGITS * 10**exp."""
    r = repr(abs(float(x)))
    mant, _, ex = r.partition("e")
    ex = int(ex) if ex else 0
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    lead = len(ip) if ip != "0" else -(len(fp) - len(fp.lstrip("0")))
    return digits.rstrip("0") or "0", lead + ex


def list_real(x):
    """One REAL(8) item of list-directed output (no separator)."""
    x = float(x)
    if x == 0.0:
        return "-0." if np.signbit(x) else "0."
    sign = "-" if x < 0 else ""
    d, e = _digits(x)
    n = len(d)
    k1 = int(f"{abs(x):.0e}".split("e")[1])   # exponent of |x| rounded to one digit
    if -1 <= k1 <= 14:                        # F form: 0.1 <= that < 1e15
        if e <= 0:
            body = "." + "0" * (-e) + d
        elif e >= n:
            body = d + "0" * (e - n) + "."
        else:
            body = d[:e] + "." + d[e:]
        return sign + body
    k = e - 1
    return f"{sign}{d[0]}.{d[1:]}E{'-' if k < 0 else '+'}{abs(k):02d}"


def list_record(values):
    """A list-directed record of REAL(8) items: each preceded by one blank."""
    return "".join(" " + list_real(v) for v in values) + "\n"


def list_int(i):
    return f" {int(i)}\n"


def e_edit(x, w, d, letter="E"):
    """Fortran Ew.d (or Dw.d) output: 0.ddd...E+ee, right-justified in w."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * d + letter + "+00"
    else:
        m, _, ex = f"{abs(x):.{d - 1}e}".partition("e")
        k = int(ex) + 1
        body = "0." + m.replace(".", "") + (f"{letter}{'-' if k < 0 else '+'}{abs(k):02d}"
                                           if abs(k) < 100 else f"{'-' if k < 0 else '+'}{abs(k):03d}")
        if x < 0:
            body = "-" + body
    return body.rjust(w)


def strain_record(v):
    """'(6(E16.8,2X))' (postprocess.f90:321, 334); the trailing blanks are not written."""
    return "  ".join(e_edit(a, 16, 8) for a in v) + "\n"


def gmres_log(hist):
    """The residual history in the reference's log format,
    FORMAT('   its =', i4, ' rel. res. norm =', d20.6) (kronecker.f90:354): one record per
    entry of ``hist`` (relative residual norms, iteration 0 first)."""
    return "".join(f"   its ={i:4d} rel. res. norm ={e_edit(r, 20, 6, 'D')}\n"
                   for i, r in enumerate(hist))


class ResultFiles:
    """The driver's result files for job ``job`` in ``directory`` (initfile +
    mpipostprocess, postprocess.f90:106-210, 380-469), one rank."""

    SUFFIX0 = ("time", "disp", "temp")
    SUFFIX1 = ("strain", "stress", "damage", "energy", "status")

    def __init__(self, job, directory=".", rank=0):
        self.job, self.dir, self.rank = job, directory, rank
        for s in self.SUFFIX0 + self.SUFFIX1:   # STATUS='unknown': created, not truncated
            open(self.path(s), "a").close()

    def path(self, suffix):
        base = self.job if suffix in self.SUFFIX0 else f"{self.job}_{self.rank}"
        return os.path.join(self.dir, f"{base}.{suffix}")

    def _append(self, suffix, text):
        with open(self.path(suffix), "a") as f:
            f.write(text)

    def write_step(self, mesh: Mesh, Nx, Ny, Nz, ym, nu, w, tcoorde, sol, tempst, gpsta,
                   elsta, D, ws, nip, nipc):
        """One output step: the same files, records and numbers as mpipostprocess."""
        i = output_index(nip, nipc)
        nt, t = temporal_functions(tcoorde, w, nip, i)
        strain, stress = gauss_strain_stress(mesh, Nx, Ny, Nz, ym, nu, sol, nt, gpsta)
        if self.rank == 0:
            self._append("time", list_record([t]))
            self._append("disp", "".join(list_record(r)
                                         for r in nodal_displacement(sol, nt, mesh.nn)))
            self._append("temp", "".join(list_record([v])
                                         for v in nodal_temperature(tempst, nt, mesh.nn)))
        self._append("strain", "".join(strain_record(r) for r in strain))
        self._append("stress", "".join(strain_record(r) for r in stress))
        self._append("damage", "".join(list_record([v]) for v in np.asarray(D)))
        self._append("energy", "".join(list_record([v]) for v in np.asarray(ws)))
        self._append("status", "".join(list_int(v) for v in np.asarray(elsta)))
        return t
