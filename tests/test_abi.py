"""The C-ABI library loads and exports every entry point include/xtfem.h declares; the
host-side parts of the product run without a GPU (no device calls here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, digest, rel

HEADER = os.path.join(ROOT, "include", "xtfem.h")
LIB = os.path.join(ROOT, "xtfem_amd", "libxtfem_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xt_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def so():
    if not os.path.exists(LIB):
        pytest.skip("libxtfem_hip.so not built")
    return C.CDLL(LIB)


def test_header_declares_the_reference_path():
    names = declared()
    for n in ("xt_kronspmv", "xt_kronspgemv", "xt_amux", "xt_ilut", "xt_lusol", "xt_kronlusol",
              "xt_kronpgmres", "xt_kokkos_host", "xt_chkelesta", "xt_slab_solve",
              "xt_damage_device", "xt_delelement", "xt_chkcnsta", "xt_chkndsta",
              "xt_adddiagonal", "xt_csrapplybc", "xt_elestatus_device"):
        assert n in names


def test_every_declared_symbol_is_exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (xt_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # the reference's own damage-boundary symbols (kokkos_host.cpp:14-41)
    for n in ("kokkos_init_", "kokkos_finish_", "kokkos_host_"):
        assert re.search(rf"\bT {n}$", out, flags=re.M), n
    for n in declared():
        getattr(so, n)


def test_python_bindings_cover_the_header():
    from xtfem_amd import _lib
    assert set(declared()) == set(_lib.EXPORTS)


def _header_params():
    """{name: [C type of each parameter]} of the functions include/xtfem.h declares."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"#.*", "", text)
    out = {}
    for m in re.finditer(r"\b(xt_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        args = " ".join(m.group(2).split())
        out[m.group(1)] = [] if args in ("", "void") else [a.strip() for a in args.split(",")]
    return out


def test_python_signatures_match_the_header():
    """Every ctypes argument list in xtfem_amd/_lib.py has the header's parameter count
    and kinds (pointers and handles as void *, int, unsigned, long long, double), so no
    call through the binding shifts its arguments."""
    import ctypes as C
    from xtfem_amd import _lib

    def kind(decl):
        if "*" in decl or re.match(r"(const\s+)?(xt_op|xt_prec|xt_asm)\b", decl):
            return C.c_void_p
        base = re.sub(r"\b(const|signed)\b", "", decl).split()
        t = " ".join(base[:-1])
        return {"int": C.c_int, "unsigned": C.c_uint, "unsigned int": C.c_uint,
                "long long": C.c_longlong, "double": C.c_double}[t]

    params = _header_params()
    assert set(params) == set(_lib._SIGS)
    for name, (_, argtypes) in _lib._SIGS.items():
        want = [kind(d) for d in params[name]]
        got = [C.c_void_p if a in (C.c_char_p, _lib.pi, _lib.pd) else a for a in argtypes]
        assert got == want, (name, params[name])


def test_fortran_interfaces_match_the_header():
    """Each BIND(C) interface of the Fortran binding has the C declaration's parameters in
    order, passed the same way: a scalar or handle by value in C has the VALUE attribute,
    a pointer parameter is passed by reference."""
    text = open(os.path.join(ROOT, "xtfem_amd", "fortran", "xtfem_hip.f90")).read()
    text = re.sub(r"&\s*\n\s*", " ", text)            # join continuation lines
    params = _header_params()
    blocks = re.findall(r"(?:function|subroutine)\s+(xt_[a-z0-9_]+)\s*\(([^)]*)\)[^\n]*"
                        r"bind\(C,\s*name=\"(xt_[a-z0-9_]+)\"\)(.*?)end (?:function|subroutine)",
                        text, flags=re.S)
    assert len(blocks) >= 15
    for fname, args, cname, body in blocks:
        assert fname == cname and cname in params, cname
        names = [a.strip() for a in args.split(",") if a.strip()]
        decls = params[cname]
        assert len(names) == len(decls), (cname, names, decls)
        byval = set()
        for line in body.split("\n"):
            if "::" in line and re.search(r"\bvalue\b", line.split("::")[0]):
                byval |= {v.strip().split("(")[0] for v in line.split("::")[1].split(",")}
        for nm, d in zip(names, decls):
            assert (nm in byval) == ("*" not in d), (cname, nm, d)


def test_tuner_knob_ids_match_the_header():
    """xtfem_amd.tune's knob names map to the XT_TUNE_* ids include/xtfem.h defines, and
    every knob the header defines is tunable from Python."""
    from xtfem_amd import tune
    text = open(HEADER).read()
    ids = {m.group(1).lower(): int(m.group(2))
           for m in re.finditer(r"#define XT_TUNE_([A-Z0-9_]+) (\d+)", text)}
    n = ids.pop("nknobs")
    assert sorted(ids.values()) == list(range(n))
    assert ids == tune.KNOBS


def test_host_ilut_bit_exact_with_oracle(oracle_lib, rod, golden):
    """xt_ilut (host C++ in the product) reproduces ilut.f: same factors, bit for bit."""
    from xtfem_amd import api
    alu, jlu, ju, ierr = api.ilut(rod["Kb"], 20, 1e-4)
    assert ierr == 0 and digest(alu, jlu, ju) == str(golden["ilut20_digest"])
    for lfil, tol in ((0, 0.0), (3, 1e-2), (50, 0.0), (12, 1e-6)):
        a = api.ilut(rod["Kb"], lfil, tol)
        b = oracle_lib.ilut(rod["Kb"], lfil, tol)
        assert a[3] == b[3] == 0
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y), (lfil, tol)


def test_host_ilut_error_codes(oracle_lib, rod):
    from xtfem_amd import api
    p = len(rod["Kb"][0]) - 1
    assert api.ilut(rod["Kb"], -1, 0.0)[3] == -4
    for iwk in (p + 10, p + 2 + 3 * p):
        assert api.ilut(rod["Kb"], 20, 1e-4, iwk=iwk)[3] == oracle_lib.ilut(rod["Kb"], 20, 1e-4, iwk=iwk)[3]
    ia = np.array([1, 3, 3, 5], dtype=np.int32)
    ja = np.array([1, 2, 2, 3], dtype=np.int32)
    a = np.array([4.0, 1.0, 1.0, 3.0])
    assert api.ilut((ia, ja, a), 2, 0.0)[3] == -5


def test_product_refuses_to_run_without_device():
    """No CPU fallback: device entry points raise when no HIP device is visible."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    from xtfem_amd import device
    with pytest.raises(Exception):
        device.KronOperator([((np.array([1, 2], np.int32)), np.array([1], np.int32), np.ones(1))])


def test_slab_matrices_vectorised_over_load_cases(oracle_lib):
    """slab.stcoeff_m broadcasts over frequencies (one set per load case)."""
    from xtfem_amd import slab
    f = np.array([10.0, 33.3, 97.0])
    c = slab.stcoeff_m(0.0, 0.5, 2 * np.pi * f)
    for r in range(3):
        o = oracle_lib.stcoeff_m(0.0, 0.5, 2 * np.pi * f[r])
        for k in o:
            assert rel(c[k][r], o[k]) < 1e-15


def test_structured_mesh_and_notch():
    from xtfem_amd import fem
    keep = np.ones((10, 6, 2), bool)
    keep[0:4, 3, :] = False
    m = fem.structured_hex(10, 6, 2, 10.0, 6.0, 2.0, keep=keep)
    assert m.ne == 120 - 8
    assert m.nn == 11 * 7 * 3                    # a one-element notch leaves no orphan node
    ke, me, Nx, Ny, Nz = fem.element_matrices(m, 2e5, 0.3, 1.0)
    # rigid translations are in the kernel of every element stiffness
    for c in range(3):
        t = np.zeros(24)
        t[c::3] = 1.0
        assert np.max(np.abs(ke @ t)) < 1e-9 * np.max(np.abs(ke))
    # total mass: each direction block of me sums to rho * volume
    vol = 10.0 * 6.0 * 2.0 * (112 / 120)
    assert abs(me[:, 0::3, 0::3].sum() - vol) < 1e-12 * vol


def test_node_orderings_are_permutations():
    """x-fastest and z-fastest numberings give the same stiffness up to a node
    permutation (matched through the coordinates)."""
    from xtfem_amd import fem
    keep = np.ones((6, 4, 2), bool)
    keep[0:2, 2, :] = False
    ma = fem.structured_hex(6, 4, 2, 6.0, 4.0, 2.0, keep=keep, order="zyx")
    mb = fem.structured_hex(6, 4, 2, 6.0, 4.0, 2.0, keep=keep, order="xyz")
    assert ma.nn == mb.nn and ma.ne == mb.ne
    key = lambda m: np.lexsort(m.xcoord.T[::-1])
    perm = np.empty(ma.nn, int)
    perm[key(ma)] = key(mb)                      # node i of a is node perm[i] of b
    assert np.array_equal(ma.xcoord, mb.xcoord[perm])
    Ka = fem.assemble(ma, 2e5, 0.3, 1.0)[0]
    Kb = fem.assemble(mb, 2e5, 0.3, 1.0)[0]
    import scipy.sparse as sp
    dense = lambda S, n: sp.csr_matrix((S[2], S[1] - 1, S[0] - 1), shape=(n, n)).toarray()
    n = 3 * ma.nn
    d = (3 * perm[:, None] + np.arange(3)).ravel()
    A, B = dense(Ka, n), dense(Kb, n)[np.ix_(d, d)]
    assert np.max(np.abs(A - B)) <= 1e-9 * np.max(np.abs(A))


def test_traction_lumping_sums_to_total_force():
    from xtfem_amd import fem
    m = fem.structured_hex(4, 3, 2, 4.0, 3.0, 2.0)
    r = fem.traction_load(m, 0, 4.0, 0, 123.0)
    assert abs(r.sum() - 123.0) < 1e-12 and np.count_nonzero(r[1::3]) == 0


FLANG = "/opt/rocm/llvm/bin/flang"


def test_fortran_binding_links_and_runs_ilut(oracle_lib, tmp_path):
    """The Fortran binding (xtfem_amd/fortran/xtfem_hip.f90) compiles, a driver calling
    the reference-named ilut links against libxtfem_hip.so, and its factors equal the
    oracle's bit for bit (ilut runs on the host, so no device is needed)."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    if not os.path.exists(LIB):
        pytest.skip("libxtfem_hip.so not built")
    fdir = os.path.join(ROOT, "xtfem_amd", "fortran")
    exe = str(tmp_path / "ilut_driver")
    libdir = os.path.dirname(LIB)
    subprocess.run([FLANG, "-c", os.path.join(fdir, "xtfem_hip.f90"), "-module-dir",
                    str(tmp_path), "-o", str(tmp_path / "xtfem_hip.o")], check=True)
    subprocess.run([FLANG, os.path.join(ROOT, "tests", "fortran", "ilut_driver.f90"),
                    "-I", str(tmp_path), str(tmp_path / "xtfem_hip.o"), "-L", libdir,
                    "-lxtfem_hip", f"-Wl,-rpath,{libdir}", "-Wl,-rpath-link,/opt/rocm/lib",
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    assert out[0].split() == ["ierr", "0"]
    n = int(out[1].split()[1])
    ia = np.array(out[3].split(), dtype=np.int32)
    ja = np.array(out[4].split(), dtype=np.int32)
    a = np.array(out[5].split(), dtype=np.float64)
    ju = np.array(out[7].split(), dtype=np.int32)
    jlu = np.array(out[8].split(), dtype=np.int32)
    alu = np.array(out[9].split(), dtype=np.float64)
    oa, ojlu, oju, oierr = oracle_lib.ilut((ia, ja, a), 3, 1e-3)
    assert oierr == 0 and len(ju) == n
    m = int(ojlu[n]) - 1
    assert np.array_equal(ju, oju) and np.array_equal(jlu, ojlu[:m])
    assert np.array_equal(alu, oa[:m])


def test_fortran_binding_live_path(tmp_path):
    """The live-path wrappers of the Fortran binding (mumps_solver_hip and mpipgmres_hip in
    place of mumps_solver and mpipgmres, the factor as a handle) compile, link against
    libxtfem_hip.so and reach it: without a device the factorisation stops with the
    library's message through xt_check, as every fatal path of the binding does; with one
    the solve of (K + 0.5 I) x = 1 (K the 1-D Laplacian of order 4) converges to it."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    if not os.path.exists(LIB):
        pytest.skip("libxtfem_hip.so not built")
    fdir = os.path.join(ROOT, "xtfem_amd", "fortran")
    exe = str(tmp_path / "live_driver")
    libdir = os.path.dirname(LIB)
    subprocess.run([FLANG, "-c", os.path.join(fdir, "xtfem_hip.f90"), "-module-dir",
                    str(tmp_path), "-o", str(tmp_path / "xtfem_hip.o")], check=True)
    subprocess.run([FLANG, os.path.join(ROOT, "tests", "fortran", "live_driver.f90"),
                    "-I", str(tmp_path), str(tmp_path / "xtfem_hip.o"), "-L", libdir,
                    "-lxtfem_hip", f"-Wl,-rpath,{libdir}", "-Wl,-rpath-link,/opt/rocm/lib",
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120).stdout
    if "failed" in out:
        assert "libxtfem_hip: mumps_solver_hip failed" in out, out
        assert "ierr" not in out, out
        return
    lines = out.split("\n")
    assert lines[0].split() == ["ierr", "0"], out
    sol = np.array(lines[1].split()[1:], dtype=np.float64)
    K = 2.0 * np.eye(4) - np.eye(4, k=1) - np.eye(4, k=-1)
    assert rel(sol, np.linalg.solve(K + 0.5 * np.eye(4), np.ones(4))) < 1e-12


def test_fortran_binding_runs_the_deletion_sequence(tmp_path):
    """delelement / chkcnsta / chkndsta / adddiagonal called by their reference names
    from Fortran (tests/fortran/delete_driver.f90) give the same CSR, counts and orphan
    list as the Python front end, bit for bit (host work, no device)."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    if not os.path.exists(LIB):
        pytest.skip("libxtfem_hip.so not built")
    from conftest import structured_system
    from xtfem_amd import api, fem
    S = structured_system(4, 2, 2, 8.0, 4.0, 4.0)
    mesh = S["mesh"]
    wqs, dns, dnxyz = fem.shape_tables_m()
    stiff = fem.elastic_dmat(2.0e5, 0.3)
    deg = fem.node_degree(mesh).astype(np.int32)
    corner = int(np.flatnonzero(deg == 1)[0]) + 1
    nef = np.unique(np.r_[np.flatnonzero((mesh.xconn == corner).any(axis=1)) + 1, 6, 11])
    nef = nef.astype(np.int32)
    K = S["Kb"]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    ndof, knz = len(K[0]) - 1, len(K[2])
    with open(fin, "wb") as f:
        f.write(np.array([mesh.nn, mesh.ne, ndof, knz, len(nef)], dtype=np.int32).tobytes())
        for a, dt in ((mesh.xconn, np.int32), (mesh.xcoord, np.float64), (dns, np.float64),
                      (dnxyz, np.float64), (fem.SD, np.float64), (stiff, np.float64)):
            f.write(np.asfortranarray(a, dtype=dt).tobytes(order="F"))
        for a in (wqs, nef, K[0], K[1], K[2], deg):
            f.write(np.ascontiguousarray(a).tobytes())
    fdir = os.path.join(ROOT, "xtfem_amd", "fortran")
    exe = str(tmp_path / "delete_driver")
    libdir = os.path.dirname(LIB)
    subprocess.run([FLANG, "-c", os.path.join(fdir, "xtfem_hip.f90"), "-module-dir",
                    str(tmp_path), "-o", str(tmp_path / "xtfem_hip.o")], check=True)
    subprocess.run([FLANG, os.path.join(ROOT, "tests", "fortran", "delete_driver.f90"),
                    "-I", str(tmp_path), str(tmp_path / "xtfem_hip.o"), "-L", libdir,
                    "-lxtfem_hip", f"-Wl,-rpath,{libdir}", "-Wl,-rpath-link,/opt/rocm/lib",
                    "-o", exe], check=True)
    subprocess.run([exe, str(fin), str(fout)], check=True)
    raw = open(fout, "rb").read()
    knz2, ndf = np.frombuffer(raw[:8], dtype=np.int32)
    off = 8
    take = lambda n, dt: np.frombuffer(raw, dtype=dt, count=n, offset=off)
    ki = take(ndof + 1, np.int32); off += 4 * (ndof + 1)
    kj = take(knz2, np.int32); off += 4 * knz2
    kv = take(knz2, np.float64); off += 8 * knz2
    necnt = take(mesh.nn, np.int32); off += 4 * mesh.nn
    ndfail = take(ndf, np.int32)
    K1 = api.delelement(mesh.xconn, mesh.xcoord, dns, dnxyz, fem.SD, stiff, wqs, nef, K)
    cnt = api.chkcnsta(mesh.nn, mesh.xconn, nef, deg)
    nd = api.chkndsta(mesh.nn, mesh.xconn, nef, cnt)
    K2 = api.adddiagonal(3, nd, K1)
    assert len(nd) >= 1 and np.array_equal(ndfail, nd) and np.array_equal(necnt, cnt)
    for a, b in zip((ki, kj, kv), K2):
        assert np.array_equal(a, b)
