"""Spatial FEM setup feeding the space-time solve: C3D8 meshes, K/M assembly, 1-0 BCs.

This is host-side input construction (it runs once per mesh), restated from the
reference's setup path so that the matrices handed to the GPU solve are the ones
the reference would build:

* Gauss points and weights: ``gaussquadrature`` n=2 (procs.f90:517-548), with the
  reference's 15-digit literal 0.577350269189626.
* C3D8 shape functions and derivatives in the reference's local node order
  (``shape3x``/``shape3dx``, procs.f90:1298-1414).
* Element matrices ``elemtx_m`` (elesubs.f90:2-58): ke = sum_l B^T D B det(J),
  me = sum_l N^T rho N det(J); D from ``setsdstiff`` (elesubs.f90:374-401);
  Nx/Ny/Nz tables (shape-function x/y/z derivatives at every Gauss point).
* Triplet -> CSR as ``tricsr`` (spmtx.f90:359-407): rows and columns sorted,
  duplicates summed in element order, exact zeros removed.
* 1-0 boundary conditions ``csrapplybc`` (spmtx.f90:268-307).
* Traction lumping per element-node incidence (main3d.f90:508-531).
* Thermal side (1 DOF a node): ``setshapefunc_t`` tables (elesubs.f90:337-371),
  conductivity/capacity matrices ``elemtx_t`` (elesubs.f90:60-116), the
  temperature-dependent material ``material_update`` (thermoelastic.f90:1-46), the
  thermal load ``vv`` that ``mpiglbmtx_m`` accumulates from ``calc_thermo_stress`` and
  ``elevec_m`` (mpisubs.f90:124-130, thermoelastic.f90:48-78, elesubs.f90:119-167), and
  the heat-flux lumping and fixed-temperature mask (main3d.f90:429-478).

All CSR output is 1-based int32 (the reference's convention); DOF numbering is
``3*(node-1) + comp`` (mpisubs.f90:111-116).
"""
from __future__ import annotations

import dataclasses
import numpy as np

GAUSS2 = np.array([-0.577350269189626, 0.577350269189626])

# local (xi, eta, zeta) sign of each C3D8 node, procs.f90:1320-1327
NODE_SIGNS = np.array([
    [-1, -1, +1], [-1, -1, -1], [-1, +1, -1], [-1, +1, +1],
    [+1, -1, +1], [+1, -1, -1], [+1, +1, -1], [+1, +1, +1]], dtype=np.float64)


@dataclasses.dataclass
class Mesh:
    xcoord: np.ndarray      # [nn, 3] float64
    xconn: np.ndarray       # [ne, 8] int32, 1-based node ids

    @property
    def nn(self):
        return self.xcoord.shape[0]

    @property
    def ne(self):
        return self.xconn.shape[0]


def read_inp(path) -> Mesh:
    """Mesh file reader (main3d.f90:216-226): nn, then nn records (id x y z), ne, then
    ne records (id n1..n8), each read as the reference's list-directed READ reads it: a
    READ starts at a new record, takes its values across records if one is short, and
    drops what is left of its last record; blanks and commas separate values."""
    with open(path) as f:
        recs = [ln.replace(",", " ").split() for ln in f]
    pos = [0]

    def read(n):
        vals = []
        while len(vals) < n:
            if pos[0] >= len(recs):
                raise ValueError(f"{path}: end of file inside a record")
            vals.extend(recs[pos[0]])
            pos[0] += 1
        return vals[:n]

    nn = int(read(1)[0])
    xc = np.zeros((nn, 3))
    for _ in range(nn):
        v = read(4)
        xc[int(v[0]) - 1] = [float(t.replace("d", "e").replace("D", "E")) for t in v[1:]]
    ne = int(read(1)[0])
    xe = np.zeros((ne, 8), dtype=np.int32)
    for _ in range(ne):
        v = read(9)
        xe[int(v[0]) - 1] = [int(t) for t in v[1:]]
    return Mesh(xc, xe)


def structured_hex(nx, ny, nz, lx, ly, lz, origin=(0.0, 0.0, 0.0), keep=None,
                   order="zyx") -> Mesh:
    """Structured C3D8 grid of nx*ny*nz elements over [0,lx]x[0,ly]x[0,lz].

    ``order`` names the node numbering from fastest to slowest axis ("zyx": z
    fastest, then y, then x; "xyz": x fastest).  ``keep`` is an optional boolean
    array [nx, ny, nz] of elements to retain (notches); orphan nodes are dropped and
    the survivors renumbered in the same order.  Local node order follows the
    reference's shape3x so det(J) > 0.
    """
    xs = origin[0] + np.linspace(0.0, lx, nx + 1)
    ys = origin[1] + np.linspace(0.0, ly, ny + 1)
    zs = origin[2] + np.linspace(0.0, lz, nz + 1)
    n1 = {"x": nx + 1, "y": ny + 1, "z": nz + 1}
    ax = {"x": 0, "y": 1, "z": 2}
    o_fast, o_mid, o_slow = order[0], order[1], order[2]

    def nid(i, j, k):
        ijk = (i, j, k)
        return (ijk[ax[o_slow]] * n1[o_mid] + ijk[ax[o_mid]]) * n1[o_fast] + ijk[ax[o_fast]]

    I0, J0, K0 = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    xcoord = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    ids = nid(I0.ravel(), J0.ravel(), K0.ravel())
    xcoord[ids, 0] = xs[I0.ravel()]
    xcoord[ids, 1] = ys[J0.ravel()]
    xcoord[ids, 2] = zs[K0.ravel()]

    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    if keep is not None:
        m = np.asarray(keep, dtype=bool).ravel()
        I, J, K = I[m], J[m], K[m]
    s = NODE_SIGNS
    conn = np.stack([nid(I + (s[a, 0] > 0), J + (s[a, 1] > 0), K + (s[a, 2] > 0))
                     for a in range(8)], axis=1)
    used = np.zeros(xcoord.shape[0], dtype=bool)
    used[conn.ravel()] = True
    newid = np.cumsum(used) - 1
    return Mesh(xcoord[used].copy(), (newid[conn] + 1).astype(np.int32))


def _shape_tables():
    """N[g, a] and dN/dxi[g, m, a] at the 8 Gauss points (l = 4i + 2j + k)."""
    pts = np.array([[GAUSS2[i], GAUSS2[j], GAUSS2[k]]
                    for i in range(2) for j in range(2) for k in range(2)])
    s = NODE_SIGNS
    f = 1.0 + s[None, :, :] * pts[:, None, :]            # [8g, 8a, 3]
    N = 0.125 * f[:, :, 0] * f[:, :, 1] * f[:, :, 2]
    dN = np.empty((8, 3, 8))
    dN[:, 0, :] = (0.125 * s[None, :, 0]) * f[:, :, 1] * f[:, :, 2]
    dN[:, 1, :] = (0.125 * s[None, :, 1]) * f[:, :, 0] * f[:, :, 2]
    dN[:, 2, :] = (0.125 * s[None, :, 2]) * f[:, :, 0] * f[:, :, 1]
    return N, dN


# Strain-displacement selector sd of setsdstiff (elesubs.f90:390-399): row = Voigt
# component (xx, yy, zz, xy, yz, xz, engineering shear), column 3*c + p holds
# d u_c / d x_p.
SD = np.zeros((6, 9))
for _r, _cols in enumerate(((0,), (4,), (8,), (1, 3), (5, 7), (2, 6))):
    SD[_r, list(_cols)] = 1.0


def shape_tables_m():
    """setshapefunc_m (elesubs.f90:299-335, shape3dx procs.f90:1344-1414) for C3D8
    with the 2-point Gauss rule: wqs [2]; dns [24, 8] with row 3l + m = dN_a/dxi_m at
    Gauss point l (l = 4i + 2j + k); dnxyz [72, 24] with dnxyz[9l + 3c + m, 3a + c] =
    dN_a/dxi_m (the derivative block of displacement component c)."""
    _, dN = _shape_tables()
    wqs = np.ones(2)
    dns = dN.reshape(24, 8).copy()
    dnxyz = np.zeros((72, 24))
    for l in range(8):
        for c in range(3):
            dnxyz[9 * l + 3 * c:9 * l + 3 * c + 3, c::3] = dN[l]
    return wqs, dns, dnxyz


def _det3(A):
    """procs.f90:476-478 (explicit 3x3 determinant, same term order)."""
    a = lambda i, j: A[..., i - 1, j - 1]
    return (a(1, 1) * a(2, 2) * a(3, 3) + a(1, 2) * a(2, 3) * a(3, 1) + a(1, 3) * a(2, 1) * a(3, 2)
            - a(1, 1) * a(2, 3) * a(3, 2) - a(1, 2) * a(2, 1) * a(3, 3) - a(1, 3) * a(2, 2) * a(3, 1))


def _inv3(A, det):
    """procs.f90:300-310 (cofactor inverse)."""
    a = lambda i, j: A[..., i - 1, j - 1]
    R = np.empty_like(A)
    R[..., 0, 0] = a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2)
    R[..., 0, 1] = a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3)
    R[..., 0, 2] = a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2)
    R[..., 1, 0] = a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3)
    R[..., 1, 1] = a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1)
    R[..., 1, 2] = a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3)
    R[..., 2, 0] = a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1)
    R[..., 2, 1] = a(1, 2) * a(3, 1) - a(1, 1) * a(3, 2)
    R[..., 2, 2] = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)
    return R / det[..., None, None]


def elastic_dmat(ym, nu):
    """setsdstiff (elesubs.f90:374-401): isotropic D in Voigt xx,yy,zz,xy,yz,xz."""
    ym = np.asarray(ym, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    c = ym / ((1.0 + nu) * (1.0 - 2.0 * nu))
    D = np.zeros(ym.shape + (6, 6))
    for i in range(3):
        for j in range(3):
            D[..., i, j] = c * nu
        D[..., i, i] = c * (1.0 - nu)
        D[..., i + 3, i + 3] = c * (0.5 - nu)
    return D


def element_matrices(mesh: Mesh, ym, nu, rho, chunk=8192):
    """elemtx_m (elesubs.f90:2-58) for every element, vectorised over elements.

    Each MATMUL of the reference is evaluated as a sequential sum in index order
    (jac = dns*X, bmat = (sd*gam)*dnxyz, ket = bmat^T*(stiff*bmat),
    met = ns^T*(rho*ns)) and ke/me accumulate over the 8 Gauss points in order, so
    the products and sums round exactly as the compiled reference does.

    ym, nu: scalars or per-Gauss-point arrays [8*ne].
    Returns ke, me [ne, 24, 24] (DOF order 3*a + comp) and Nx, Ny, Nz [ne*8, 8]
    (row g = Gauss point, column a = local node).
    """
    ne = mesh.ne
    N, dN = _shape_tables()
    ymg = np.broadcast_to(np.asarray(ym, dtype=np.float64), (ne * 8,)).reshape(ne, 8)
    nug = np.broadcast_to(np.asarray(nu, dtype=np.float64), (ne * 8,)).reshape(ne, 8)
    ke = np.zeros((ne, 24, 24))
    me = np.zeros((ne, 24, 24))
    Gall = np.zeros((ne, 8, 3, 8))
    for e0 in range(0, ne, chunk):
        e1 = min(ne, e0 + chunk)
        X = mesh.xcoord[mesh.xconn[e0:e1] - 1]             # [E, 8a, 3p]
        E = e1 - e0
        for l in range(8):
            jac = np.zeros((E, 3, 3))
            for a in range(8):                              # jac(m,p) = sum_a dns(m,a) X(a,p)
                jac = jac + dN[l, :, a][None, :, None] * X[:, a, None, :]
            det = _det3(jac)
            if np.any(det <= 0.0):
                raise ValueError("non-positive Jacobian: element node order does not match C3D8")
            gam = _inv3(jac, det)
            G = np.zeros((E, 3, 8))                         # dN_a/dx_p = sum_m gam(p,m) dN(m,a)
            for m in range(3):
                G = G + gam[:, :, m, None] * dN[l, m, :][None, None, :]
            Gall[e0:e1, l] = G
            B = np.zeros((E, 6, 24))
            gx, gy, gz = G[:, 0, :], G[:, 1, :], G[:, 2, :]
            B[:, 0, 0::3] = gx
            B[:, 1, 1::3] = gy
            B[:, 2, 2::3] = gz
            B[:, 3, 0::3] = gy
            B[:, 3, 1::3] = gx
            B[:, 4, 1::3] = gz
            B[:, 4, 2::3] = gy
            B[:, 5, 0::3] = gz
            B[:, 5, 2::3] = gx
            D = elastic_dmat(ymg[e0:e1, l], nug[e0:e1, l])  # [E, 6, 6]
            DB = np.zeros((E, 6, 24))
            for j in range(6):
                DB = DB + D[:, :, j, None] * B[:, j, None, :]
            ket = np.zeros((E, 24, 24))
            for i in range(6):
                ket = ket + B[:, i, :, None] * DB[:, i, None, :]
            ke[e0:e1] = ke[e0:e1] + ket * det[:, None, None]
            mm = N[l, :, None] * (rho * N[l, None, :])      # [8, 8]
            met = np.zeros((24, 24))
            for c in range(3):
                met[c::3, c::3] = mm
            me[e0:e1] = me[e0:e1] + met[None] * det[:, None, None]
    Nx = Gall[:, :, 0, :].reshape(ne * 8, 8).copy()
    Ny = Gall[:, :, 1, :].reshape(ne * 8, 8).copy()
    Nz = Gall[:, :, 2, :].reshape(ne * 8, 8).copy()
    return ke, me, Nx, Ny, Nz


def _triplet_to_csr(rows, cols, vals, n):
    """tricsr: sort by (row, col), sum duplicates sequentially in input (element)
    order, drop exact zeros.  rows/cols 0-based in, 1-based CSR out."""
    key = rows.astype(np.int64) * n + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    start = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    cnt = np.diff(np.r_[start, len(key)])
    acc = vals[start].copy()
    for k in range(1, int(cnt.max()) if len(cnt) else 1):
        m = cnt > k
        acc[m] = acc[m] + vals[start[m] + k]
    ukey = key[start]
    nz = acc != 0.0
    ukey, acc = ukey[nz], acc[nz]
    r = (ukey // n).astype(np.int64)
    c = (ukey % n).astype(np.int32)
    ia = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ia, r + 1, 1)
    ia = np.cumsum(ia)
    return (ia + 1).astype(np.int32), (c + 1).astype(np.int32), acc


def assemble(mesh: Mesh, ym, nu, rho):
    """Global K, M (1-based CSR) and Nx/Ny/Nz as mpiglbmtx_m + tricsr build them."""
    ke, me, Nx, Ny, Nz = element_matrices(mesh, ym, nu, rho)
    ne = mesh.ne
    n = 3 * mesh.nn
    dofs = (3 * (mesh.xconn.astype(np.int64) - 1)[:, :, None] + np.arange(3)[None, None, :]).reshape(ne, 24)
    rows = np.repeat(dofs[:, :, None], 24, axis=2).ravel()
    cols = np.repeat(dofs[:, None, :], 24, axis=1).ravel()
    K = _triplet_to_csr(rows, cols, ke.ravel(), n)
    M = _triplet_to_csr(rows, cols, me.ravel(), n)
    return K, M, Nx, Ny, Nz


def applybc(S, bc):
    """csrapplybc (spmtx.f90:268-307): 1-0 method on rows/cols with bc != 0."""
    ia, ja, a = S
    n = len(ia) - 1
    bc = np.asarray(bc) != 0
    row = np.repeat(np.arange(n), np.diff(ia))
    col = ja - 1
    a = a.copy()
    hit = bc[row] | bc[col]
    a[hit] = 0.0
    a[hit & (row == col)] = 1.0
    keep = a != 0.0
    cnt = np.bincount(row[keep], minlength=n)
    ia2 = np.r_[0, np.cumsum(cnt)].astype(np.int32) + 1
    return ia2, ja[keep].copy(), a[keep].copy()


def face_nodes(mesh: Mesh, axis, value, tol=1e-9):
    return np.flatnonzero(np.abs(mesh.xcoord[:, axis] - value) <= tol)


def traction_load(mesh: Mesh, axis, value, direction, total_force, tol=1e-9):
    """Lumped traction as main3d.f90:508-531: each element-node incidence on the
    loaded face receives total_force / (number of such incidences)."""
    on = np.abs(mesh.xcoord[:, axis] - value) <= tol
    inc = on[mesh.xconn - 1]
    nt = inc.sum()
    rext = np.zeros(3 * mesh.nn)
    if nt == 0:
        return rext
    fnode = total_force / float(nt)
    nodes = (mesh.xconn - 1)[inc]
    np.add.at(rext, 3 * nodes + direction, fnode)
    return rext


def clamp_mask(mesh: Mesh, axis, values, tol=1e-9):
    """Spatial DOF mask (length 3*nn) fixing all 3 components of nodes on the
    planes x_axis == v for v in values (main3d.f90:482-488)."""
    m = np.zeros(mesh.nn, dtype=bool)
    for v in np.atleast_1d(values):
        m |= np.abs(mesh.xcoord[:, axis] - v) <= tol
    return np.repeat(m, 3).astype(np.int32)


def node_adjacency(mesh: Mesh):
    """Node graph of the mesh (nodes sharing an element) as CSR (0-based)."""
    c = mesh.xconn.astype(np.int64) - 1
    rows = np.repeat(c, 8, axis=1).ravel()
    cols = np.tile(c, (1, 8)).ravel()
    key = np.unique(rows * mesh.nn + cols)
    r, cc = key // mesh.nn, key % mesh.nn
    keep = r != cc
    r, cc = r[keep], cc[keep]
    ptr = np.zeros(mesh.nn + 1, dtype=np.int64)
    np.add.at(ptr, r + 1, 1)
    return np.cumsum(ptr), cc


def dof_pattern(mesh: Mesh, components="all"):
    """1-based CSR pattern (rowptr, colidx; int32) of the 3-DOF matrices of the mesh with
    every coupling its elements make, stored zeros included: all 3 x 3 component pairs of
    nodes sharing an element (components "all": K) or equal components only ("equal":
    M), columns sorted.  The assembled matrices (before the 1-0 method) and every matrix
    an element deletion leaves have their entries in it."""
    ptr, adj = node_adjacency(mesh)
    nn = mesh.nn
    deg = np.diff(ptr) + 1                          # neighbours and the node itself
    allv = np.concatenate([adj, np.arange(nn)])
    allr = np.concatenate([np.repeat(np.arange(nn), np.diff(ptr)), np.arange(nn)])
    nodes = allv[np.lexsort((allv, allr))]          # each node row sorted by node
    off = np.concatenate([[0], np.cumsum(deg)])     # node row n: nodes[off[n]:off[n+1]]
    idx = np.repeat(np.arange(nn), 3)               # node of DOF row 3 n + a
    if components == "all":
        # DOF row 3 n + a: the columns 3 m + b for the nodes m of row n, b = 0, 1, 2
        cols_node = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        lens = 3 * deg[idx]
        starts = 3 * off[idx]
    else:
        # DOF row 3 n + a: the columns 3 m + a
        lens = deg[idx]
        starts = off[idx]
    rp = np.zeros(3 * nn + 1, dtype=np.int64)
    rp[1:] = np.cumsum(lens)
    g = np.repeat(starts - rp[:-1], lens) + np.arange(rp[-1])
    if components == "all":
        cols = cols_node[g]
    else:
        cols = 3 * nodes[g] + np.repeat(np.tile(np.arange(3), nn), lens)
    return (rp + 1).astype(np.int32), (cols + 1).astype(np.int32)


def rcm_order(mesh: Mesh):
    """Reverse Cuthill-McKee order of the nodes: perm[new] = old.  Breadth-first from a
    pseudo-peripheral node of each connected part (lowest degree, then the farthest node
    of a first sweep), neighbours visited by increasing degree, the sequence reversed.
    The spatial matrices of the renumbered mesh are banded with about the cross-section's
    node count as half-bandwidth, which the banded exact factor needs; the shipped
    unstructured meshes are numbered without regard to bandwidth."""
    ptr, adj = node_adjacency(mesh)
    deg = np.diff(ptr)
    nn = mesh.nn
    visited = np.zeros(nn, dtype=bool)
    order = []

    def bfs(start):
        seq = [start]
        seen = {start}
        head = 0
        while head < len(seq):
            v = seq[head]
            head += 1
            nb = adj[ptr[v]:ptr[v + 1]]
            nb = nb[~visited[nb]]
            nb = [u for u in nb[np.argsort(deg[nb], kind="stable")] if u not in seen]
            seen.update(nb)
            seq.extend(nb)
        return seq

    for root in np.argsort(deg, kind="stable"):
        if visited[root]:
            continue
        seq = bfs(int(root))               # a first sweep finds a far node
        seq = bfs(int(seq[-1]))
        visited[seq] = True
        order.extend(seq)
    return np.asarray(order[::-1], dtype=np.int64)


def renumber(mesh: Mesh, perm):
    """The mesh with node perm[i] renamed i + 1 (element order and local node order
    unchanged, so every assembled sum keeps its order: K' = P K P^T exactly)."""
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty(mesh.nn, dtype=np.int64)
    inv[perm] = np.arange(mesh.nn)
    return Mesh(mesh.xcoord[perm].copy(), (inv[mesh.xconn - 1] + 1).astype(np.int32))


def node_bandwidth(mesh: Mesh):
    """Largest node-id distance inside an element (DOF half-bandwidth = 3 x this + 2)."""
    c = mesh.xconn
    return int((c.max(axis=1) - c.min(axis=1)).max())


def node_degree(mesh: Mesh):
    """Number of elements sharing each node (main3d.f90:273-280)."""
    return np.bincount((mesh.xconn - 1).ravel(), minlength=mesh.nn)


# ---------------------------------------------------------------------------------
# Thermal side
# ---------------------------------------------------------------------------------

# Default-real (FP32) literals of material_update (thermoelastic.f90:34-35) and of the
# driver's thermal expansion coefficient (main3d.f90:157), widened to FP64 as the
# compiled reference widens them.
YM_DT1 = float(np.float32(1.6))
YM_DT2 = float(np.float32(0.1))
NU_DT1 = float(np.float32(5.0e-5))
ALPHA_REF = float(np.float32(1.2e-5))


def shape_tables_t():
    """setshapefunc_t (elesubs.f90:337-371; shapex procs.f90:1416-1448, shapedx
    procs.f90:1452-1501): wqs [2], ns [8 GP, 8 nodes], dns [8 GP, 8 nodes, 3]."""
    N, dN = _shape_tables()
    return np.ones(2), N.copy(), np.ascontiguousarray(np.transpose(dN, (0, 2, 1)))


def _thermal_geometry(X, dN_l):
    """Jacobian, determinant and shape-function gradients at one Gauss point the way
    elemtx_t and elevec_m form them (elesubs.f90:86-95, 143-152): jac(m, p) =
    sum_a dns(a, p) X(a, m) -- the transpose of elemtx_m's jac, so the determinant and
    cofactors round differently -- then Niix(m) = sum_p gam(p, m) dns(a, p).
    X [E, 8a, 3], dN_l [3, 8a]; returns det [E] and G [E, 3m, 8a]."""
    E = X.shape[0]
    jac = np.zeros((E, 3, 3))
    for a in range(8):
        jac = jac + X[:, a, :, None] * dN_l[None, None, :, a]
    det = _det3(jac)
    if np.any(det <= 0.0):
        raise ValueError("non-positive Jacobian: element node order does not match C3D8")
    gam = _inv3(jac, det)
    G = np.zeros((E, 3, 8))
    for p in range(3):
        G = G + gam[:, p, :, None] * dN_l[p][None, None, :]
    return det, G


def thermal_element_matrices(mesh: Mesh, kappa, cp, rho, chunk=8192):
    """elemtx_t (elesubs.f90:60-116) for every element: ke [ne, 8, 8] =
    sum_l sum_m kappa Niix(m) Njjx(m) det, me = sum_l rho cp N_i N_j det, accumulated
    Gauss point by Gauss point and, inside ke, direction by direction, with the
    reference's left-to-right products (the weights are 1 for the 2-point rule)."""
    ne = mesh.ne
    N, dN = _shape_tables()
    ke = np.zeros((ne, 8, 8))
    me = np.zeros((ne, 8, 8))
    rc = rho * cp
    for e0 in range(0, ne, chunk):
        e1 = min(ne, e0 + chunk)
        X = mesh.xcoord[mesh.xconn[e0:e1] - 1]
        k = np.zeros((e1 - e0, 8, 8))
        m = np.zeros((e1 - e0, 8, 8))
        for l in range(8):
            det, G = _thermal_geometry(X, dN[l])
            mm = (rc * N[l])[:, None] * N[l][None, :]
            m = m + mm[None] * det[:, None, None]
            for d in range(3):
                k = k + ((kappa * G[:, d, :, None]) * G[:, d, None, :]) * det[:, None, None]
        ke[e0:e1] = k
        me[e0:e1] = m
    return ke, me


def assemble_t(mesh: Mesh, kappa, cp, rho):
    """Thermal K_t, M_t (1-based CSR over nodes) as mpiglbmtx_t (mpisubs.f90:141-188)
    + tricsr build them."""
    ke, me = thermal_element_matrices(mesh, kappa, cp, rho)
    ne, n = mesh.ne, mesh.nn
    dofs = (mesh.xconn.astype(np.int64) - 1)
    rows = np.repeat(dofs[:, :, None], 8, axis=2).ravel()
    cols = np.repeat(dofs[:, None, :], 8, axis=1).ravel()
    return _triplet_to_csr(rows, cols, ke.ravel(), n), _triplet_to_csr(rows, cols, me.ravel(), n)


def gauss_temperature(mesh: Mesh, T):
    """Temperature at every Gauss point, sum_a N_a T(node_a) in local node order
    (thermoelastic.f90:26-30 and 65-69 form it identically).  T [nn] -> [ne, 8]."""
    N, _ = _shape_tables()
    Te = np.asarray(T, dtype=np.float64)[mesh.xconn - 1]          # [ne, 8a]
    t = np.zeros((mesh.ne, 8))
    for a in range(8):
        t = t + N[None, :, a] * Te[:, a, None]
    return t


def material_update(mesh: Mesh, T, Tref, ym0, nu0):
    """thermoelastic.f90:1-46: E = E0 - 1.6 (T - Tref) - 0.1 (T - Tref)^2 and
    nu = nu0 + 5e-5 (T - Tref) at Gauss points whose temperature is >= 0, E0 and nu0
    elsewhere.  T: the nodal temperatures of the slab's third temporal node [nn].
    Returns ym, nu [8*ne]."""
    t = gauss_temperature(mesh, T)
    d = t - Tref
    ok = t >= 0.0
    ym = np.where(ok, (ym0 - YM_DT1 * d) - (YM_DT2 * d) * d, ym0)
    nu = np.where(ok, nu0 + NU_DT1 * d, nu0)
    return ym.ravel(), nu.ravel()


def thermal_load(mesh: Mesh, T, Tref, alpha, ym, nu, chunk=8192):
    """The thermal load vv [3*nn] that mpiglbmtx_m accumulates (mpisubs.f90:124-130):
    per Gauss point the thermal stress sigma_m = (D(m,1) + D(m,2) + D(m,3)) alpha (T -
    Tref) (calc_thermo_stress, thermoelastic.f90:48-78; D from setsdstiff with that
    point's E and nu), then ve(3a + m) += dN_a/dx_m sigma_m det (elevec_m,
    elesubs.f90:119-167, geometry as elemtx_t), summed into vv element by element."""
    ne = mesh.ne
    t = gauss_temperature(mesh, T)
    strain = alpha * (t - Tref)                                    # [ne, 8]
    D = elastic_dmat(np.asarray(ym, dtype=np.float64).reshape(ne, 8),
                     np.asarray(nu, dtype=np.float64).reshape(ne, 8))
    rs = (D[..., :3, 0] + D[..., :3, 1]) + D[..., :3, 2]           # [ne, 8, 3m]
    stress = rs * strain[..., None]
    _, dN = _shape_tables()
    ve = np.zeros((ne, 8, 3))
    for e0 in range(0, ne, chunk):
        e1 = min(ne, e0 + chunk)
        X = mesh.xcoord[mesh.xconn[e0:e1] - 1]
        v = np.zeros((e1 - e0, 8, 3))
        for l in range(8):
            det, G = _thermal_geometry(X, dN[l])
            v = v + (np.transpose(G, (0, 2, 1)) * stress[e0:e1, l, None, :]) * det[:, None, None]
        ve[e0:e1] = v
    vv = np.zeros(3 * mesh.nn)
    dofs = 3 * (mesh.xconn.astype(np.int64) - 1)[:, :, None] + np.arange(3)[None, None, :]
    np.add.at(vv, dofs.ravel(), ve.ravel())       # sequential: element order per DOF
    return vv


def fixed_temperature_mask(mesh: Mesh, axis, values, tol=0.0):
    """Spatial thermal DOF mask [nn] (one DOF a node) fixing the temperature on the
    planes x_axis == v (main3d.f90:432-440; the reference compares exactly)."""
    m = np.zeros(mesh.nn, dtype=bool)
    for v in np.atleast_1d(values):
        m |= np.abs(mesh.xcoord[:, axis] - v) <= tol
    return m.astype(np.int32)


def heat_flux_load(mesh: Mesh, axis, count_value, values, total, tol=0.0):
    """Lumped heat flux (main3d.f90:455-478): fnode = total / (element-node incidences
    on the plane x_axis == count_value); every incidence on a plane in ``values`` then
    receives fnode (the reference counts one face and loads two)."""
    xa = mesh.xcoord[:, axis]
    inc_count = (np.abs(xa - count_value) <= tol)[mesh.xconn - 1].sum()
    r = np.zeros(mesh.nn)
    if inc_count == 0:
        return r
    fnode = total / float(inc_count)
    on = np.zeros(mesh.nn, dtype=bool)
    for v in np.atleast_1d(values):
        on |= np.abs(xa - v) <= tol
    nodes = (mesh.xconn - 1)[on[mesh.xconn - 1]]
    np.add.at(r, nodes, fnode)
    return r
