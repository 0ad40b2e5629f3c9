"""The reference's mesh and restart files (oversetdevfort_amd/refio.py), both
ways, against the reference's own routines (fixture written by the
reference harness, mode 7, with this build's writers):

* this build's write_part / write_l2b / write_bmesh / write_restart give,
  byte for byte, the files the fixture holds, and what the reference's
  input_mesh, input_bmesh and readSol read from those files equals the
  arrays written, bit for bit;
* this build's read_restart recovers, bit for bit, the fields the
  reference's writeSol was given (its file's text is in the fixture; writeSol
  prints the old temperature and rate, which the harness set);
* read_part / read_l2b / read_bmesh recover what was written.
"""
import os

import numpy as np
import pytest

from conftest import golden
from oversetdevfort_amd import refio as R
from oversetdevfort_amd.mesh import Mesh


@pytest.fixture(scope="module")
def z():
    return golden("io_files")


def _bounds(z):
    out = []
    for i in range(2):
        out.append({k: z[f"bound_{k}{i}"] for k in
                    ("face_id", "face_ien", "f2e", "face_or", "bnodes")})
    return out


def _mesh(z):
    return Mesh(xg=z["xg"], ien=z["ien"], elm_id=z["elm_id"])


def test_writers_give_the_files_the_reference_read(z, tmp_path):
    m, b = _mesh(z), _bounds(z)
    R.write_part(tmp_path / R.part_name(1), m, z["node_id"], b)
    R.write_l2b(tmp_path / R.l2b_name(1), b)
    R.write_bmesh(tmp_path / R.bmesh_name(11), m, b[0], 1)
    R.write_bmesh(tmp_path / R.bmesh_name(12), m, b[1], 2)
    fields = {k: z["in_" + k] for k in R.RESTART_FIELDS}
    R.write_restart(tmp_path / R.restart_name(7), 7, float(z["t_read"]), fields)
    for key, name in (("part_text", R.part_name(1)), ("l2b_text", R.l2b_name(1)),
                      ("bmesh0_text", R.bmesh_name(11)), ("bmesh1_text", R.bmesh_name(12)),
                      ("restart_text", R.restart_name(7))):
        assert (tmp_path / name).read_text() == str(z[key]), key


def test_reference_reads_what_was_written(z):
    m, b = _mesh(z), _bounds(z)
    assert np.array_equal(z["ref_xg"], m.xg)
    assert np.array_equal(z["ref_node_id"], z["node_id"])
    assert np.array_equal(z["ref_ien"], m.ien)
    assert np.array_equal(z["ref_elm_id"], m.elm_id)
    assert np.all(z["ref_elmnshl"] == 8)
    for i in range(2):
        for k in ("face_id", "face_ien", "f2e", "face_or", "bnodes"):
            assert np.array_equal(z[f"ref_b{i}_{k}"], b[i][k]), (i, k)
        assert np.all(z[f"ref_b{i}_nshlb"] == 4)
        assert np.array_equal(z[f"ref_b{i}_l2snode"], np.arange(len(b[i]["bnodes"])))
        assert np.array_equal(z[f"ref_b{i}_l2selem"], np.arange(len(b[i]["f2e"])))
        assert np.array_equal(z[f"ref_bmesh{i}_xg"], m.xg[b[i]["bnodes"]])
        assert np.all(z[f"ref_bmesh{i}_owned"] == 1)
    assert int(z["ref_rstep"]) == 7 and float(z["ref_time"]) == float(z["t_read"])
    for k in R.RESTART_FIELDS:
        assert np.array_equal(z["ref_read_" + k], z["in_" + k]), k


def test_reader_recovers_the_reference_restart(z, tmp_path):
    path = tmp_path / R.restart_name(9)
    path.write_text(str(z["ref_restart_text"]))
    got = R.read_restart(str(path), len(z["xg"]))
    assert got["step"] == 9 and got["time"] == float(z["t_write"])
    for k in R.RESTART_FIELDS:
        assert np.array_equal(got[k], z["w_" + k]), k


def test_readers_recover_what_was_written(z, tmp_path):
    m, b = _mesh(z), _bounds(z)
    (tmp_path / "p.dat").write_text(str(z["part_text"]))
    got = R.read_part(str(tmp_path / "p.dat"))
    assert np.array_equal(got["mesh"].xg, m.xg) and np.array_equal(got["mesh"].ien, m.ien)
    assert np.array_equal(got["mesh"].elm_id, m.elm_id)
    assert np.array_equal(got["node_id"], z["node_id"])
    for i in range(2):
        for k in ("face_id", "face_ien", "f2e", "face_or", "bnodes"):
            assert np.array_equal(np.asarray(got["bounds"][i][k]), b[i][k]), (i, k)
    (tmp_path / "l.dat").write_text(str(z["l2b_text"]))
    l2b = R.read_l2b(str(tmp_path / "l.dat"), 2)
    for i in range(2):
        assert np.array_equal(l2b[i]["vnodes"], b[i]["bnodes"])
        assert np.array_equal(l2b[i]["f2e"], b[i]["f2e"])
    (tmp_path / "b.dat").write_text(str(z["bmesh1_text"]))
    bm = R.read_bmesh(str(tmp_path / "b.dat"))
    assert bm["face_id"] == 12 and np.array_equal(bm["xg"], m.xg[b[1]["bnodes"]])
    local = {int(n): j for j, n in enumerate(b[1]["bnodes"])}
    assert np.array_equal(bm["face_ien"],
                          np.vectorize(local.get)(b[1]["face_ien"]))
    (tmp_path / "r.dat").write_text(str(z["restart_text"]))
    rs = R.read_restart(str(tmp_path / "r.dat"), len(z["xg"]))
    for k in R.RESTART_FIELDS:
        assert np.array_equal(rs[k], z["in_" + k]), k


def test_list_directed_forms():
    """r*c repeats, D exponents, values across records, a short record."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, R.restart_name(1))
        with open(p, "w") as f:
            f.write("1 2.5D-01\n" + "3*1.5\n" * 2 + "0.0 1.0\n2.0\n" +
                    "\n".join(["1.0D0"] * 2 * 1) + "\n" + "4*0\n" * 2 + "\n".join(["7"] * 30) + "\n")
        got = R.read_restart(p, 2)
        assert got["step"] == 1 and got["time"] == 0.25
        assert np.array_equal(got["dg"], np.full((2, 3), 1.5))
        assert np.array_equal(got["ug"], np.array([[0.0, 1.0, 2.0], [1.0, 1.0, 0.0]]))
