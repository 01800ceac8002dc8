"""Result files and logs in the reference's formats (xtfem_amd.postprocess, fem.read_inp)
against the compiled reference's own output (tests/golden: mpipostprocess run on a
distorted block; kronpgmres's gmres.log records), plus what the mathematics fixes
(a uniform strain field is recovered exactly, the stress law)."""
import numpy as np
import pytest

from conftest import RHO_REF, rel


def _pp_inputs(golden):
    from xtfem_amd import fem
    mesh = fem.Mesh(golden["pp_xcoord"], golden["pp_xconn"])
    _, _, Nx, Ny, Nz = fem.assemble(mesh, golden["pp_ym"], golden["pp_nu"], RHO_REF)
    return mesh, Nx, Ny, Nz


def _write(tmp_path, golden):
    from xtfem_amd import postprocess as pp
    mesh, Nx, Ny, Nz = _pp_inputs(golden)
    rf = pp.ResultFiles("job", str(tmp_path))
    rf.write_step(mesh, Nx, Ny, Nz, golden["pp_ym"], golden["pp_nu"], 2 * np.pi * 20,
                  golden["pp_tc"], golden["pp_sol"], golden["pp_temp"], golden["pp_gpsta"],
                  golden["pp_elsta"], golden["pp_D"], golden["pp_ws"], 2560, 256)
    out = {}
    for p in tmp_path.iterdir():
        out[p.name.rsplit(".", 1)[1]] = p.read_text()
    return out


def _form(tok):
    """Layout class of a list-directed item: sign, F or E form, exponent width."""
    t = tok.lstrip("-")
    if "E" in t:
        return (tok.startswith("-"), "E", len(t.split("E")[1]))
    return (tok.startswith("-"), "F", t.startswith("."))


def test_edit_descriptor_files_byte_identical(tmp_path, golden):
    """strain/stress ('(6(E16.8,2X))'), status (integers) and time: the same bytes."""
    mine = _write(tmp_path, golden)
    for k in ("strain", "stress", "status", "time"):
        assert mine[k] == str(golden[f"pp_file_{k}"]), k


def test_list_directed_files_same_numbers_and_layout(tmp_path, golden):
    """disp, temp, damage, energy: same records, same item layout, and every item reads
    back as the same double (the last digit of a 17-digit item is a runtime choice)."""
    mine = _write(tmp_path, golden)
    for k in ("disp", "temp", "damage", "energy"):
        ref_lines = str(golden[f"pp_file_{k}"]).splitlines()
        my_lines = mine[k].splitlines()
        assert len(ref_lines) == len(my_lines), k
        for a, b in zip(ref_lines, my_lines):
            ta, tb = a.split(), b.split()
            assert len(ta) == len(tb)
            assert [float(x) for x in ta] == [float(x) for x in tb], (k, a, b)
            assert [_form(x) for x in ta] == [_form(x) for x in tb], (k, a, b)


def test_files_append_and_are_created(tmp_path, golden):
    from xtfem_amd import postprocess as pp
    _write(tmp_path, golden)
    first = (tmp_path / "job_0.status").read_text()
    pp.ResultFiles("job", str(tmp_path))          # initfile does not truncate
    assert (tmp_path / "job_0.status").read_text() == first
    again = _write(tmp_path, golden)
    assert again["status"] == first + first
    assert sorted(p.name for p in tmp_path.iterdir()) == sorted(
        ["job.time", "job.disp", "job.temp"] +
        [f"job_0.{s}" for s in ("strain", "stress", "damage", "energy", "status")])


def test_gauss_strain_recovers_uniform_field(golden):
    """u = G x (a linear field, the same at every temporal function) gives strain G's
    symmetric part at every Gauss point, whatever the element distortion; stress =
    D strain with D = setsdstiff's matrix."""
    from xtfem_amd import fem, postprocess as pp
    mesh, Nx, Ny, Nz = _pp_inputs(golden)
    G = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [0.0, 7e-4, 5e-4]])
    u = mesh.xcoord @ G.T
    nt = np.array([0.3, 0.5, 0.2, 0.0, 0.0, 0.0])            # sums to 1
    sol = np.tile(u.ravel(), 6)
    e, s = pp.gauss_strain_stress(mesh, Nx, Ny, Nz, golden["pp_ym"], golden["pp_nu"], sol, nt)
    want = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1],
                     G[0, 2] + G[2, 0]])
    assert rel(e, np.tile(want, (len(e), 1))) < 1e-11
    D = fem.elastic_dmat(golden["pp_ym"], golden["pp_nu"])
    assert rel(s, np.einsum("gij,gj->gi", D, e)) < 1e-12


def test_output_point_and_time():
    from xtfem_amd import postprocess as pp, slab
    i = pp.output_index(2560, 256)
    assert i == 2560 - 64 * 3 + 1
    nt, t = pp.temporal_functions([0.5, 0.75, 1.0], 2 * np.pi * 20, 2560, i)
    assert abs(t - (0.5 + (i - 1) * 0.5 / 2560)) < 1e-15
    ref = slab.damage_nt([0.5, 0.75, 1.0], 2 * np.pi * 20, 2560)[i - 1]
    assert np.allclose(nt, ref, rtol=0, atol=1e-15)
    assert abs(nt[:3].sum() - 1.0) < 1e-14


def test_gmres_log_matches_reference(oracle_lib, rod, golden):
    """The oracle's residual history (itself the reference's iterates, bit for bit) in
    the log format gives the reference's gmres.log records byte for byte."""
    from xtfem_amd import postprocess as pp
    alu, jlu, ju, _ = oracle_lib.ilut(rod["Kb"], 20, 1e-4)
    for tag, eps in (("e8", 1e-8), ("e12", 1e-12)):
        _, ierr, its, hist = oracle_lib.kronpgmres(golden["A"], golden["B"], rod["Kb"], rod["Mb"],
                                                   (alu, jlu, ju), golden["gmres_rhs"], im=30,
                                                   eps=eps, maxits=3000)
        assert ierr == 0
        assert pp.gmres_log(hist) == str(golden[f"gmres_{tag}_log"])


@pytest.mark.parametrize("x,s", [(0.0, "0."), (-0.0, "-0."), (1.0, "1."), (0.4625, ".4625"),
                                 (-2.5, "-2.5"), (1e-300, "1.E-300"), (197000.0, "197000."),
                                 (0.0246, "2.46E-02"), (15.5, "15.5"), (123.0, "123."),
                                 (0.095, ".095"), (0.0949999, "9.49999E-02"),
                                 (9.49e14, "949000000000000."), (9.5e14, "9.5E+14"),
                                 (1e15, "1.E+15"), (5e-324, "5.E-324"), (1e100, "1.E+100"),
                                 (2.5e-5, "2.5E-05")])
def test_list_directed_items(x, s):
    from xtfem_amd import postprocess as pp
    assert pp.list_real(x) == s


def test_read_inp_records(tmp_path, rod):
    """List-directed record semantics: extra values on a record are dropped, a short
    record continues on the next, commas separate."""
    from xtfem_amd import fem
    p = tmp_path / "m.inp"
    p.write_text("2 extra\n1, 0.0, 0.0, 0.0 trailing\n2 1.5\n 2.5 3.5\n1\n1 1 2 1 2 1 2 1 2 9 9\n")
    m = fem.read_inp(str(p))
    assert m.nn == 2 and m.ne == 1
    assert np.array_equal(m.xcoord, [[0.0, 0.0, 0.0], [1.5, 2.5, 3.5]])
    assert np.array_equal(m.xconn, [[1, 2, 1, 2, 1, 2, 1, 2]])
