"""The oracle's boundary-face assembly (FaceAssembly_NS_weak followed by
FaceAssembly_NS_outflow, as assembleNavSto calls them), CPU:

* tet faces against the reference's own routines (relinked harness, mode 6):
  every LHS array and RHS bit for bit, with 3-point and 1-point faces, one
  and both subdomains, and the vector-only call;
* hex faces (no reference code; the face map is an extension) by closed
  forms that a wrong normal, area Jacobian or face map would break:
  - the pressure-gradient face block G and the continuity block D1 integrate
    N_a N_b n over the boundary, so sum_ij G_ij x_j,d = fact2 * volume (the
    divergence theorem for x_d e_d) and D1 = -G;
  - the continuity RHS is the flux of u through the boundary: zero for a
    constant u, 3 * volume for u = x (which also fixes the normal outward);
  - the residual vanishes when u equals the wall (mesh) velocity, is
    uniform and p = 0.
"""
import numpy as np
import pytest

from conftest import golden
from oversetdevfort_amd import mesh as M
from oversetdevfort_amd.mesh import Mesh

PRM = np.array([0.8333333333333334, 0.6666666666666666, 0.6666666666666667,
                0.34722222222222227, 0.01, 1.3, 0.02, 0.0, 0.1, -0.2, -9.81])
FACT2 = PRM[1] * PRM[2] * PRM[4]


@pytest.mark.parametrize("name", ["face_tet_a", "face_tet_b", "face_tet_vec"])
def test_tet_faces_match_reference(oracle_mod, name):
    O = oracle_mod
    z = golden(name)
    m = Mesh(xg=z["xg"], ien=z["ien"], elm_id=z["elm_id"])
    col, row = O.sparsity(m.nnode, m.ien)
    assert np.array_equal(col, z["col"]) and np.array_equal(row, z["row"])
    f2e, fo = M.boundary_faces(m)
    assert np.array_equal(f2e, z["f2e"]) and np.array_equal(fo, z["face_or"])
    st = {k[3:]: z[k] for k in z.files if k.startswith("st_")}
    R = O.face_ns(m, z["ibc"], st, z["prm"], col, row, (f2e, fo), int(z["tensor_flag"]),
                  int(z["subdomain_flag"]), ngb=int(z["ngb"]))
    for k in ("k11", "g", "d1", "m", "rhsu", "rhsp"):
        assert np.array_equal(R[k], z[k]), k
    assert np.abs(z["rhsu"]).max() > 0.1           # the faces are not empty


def _box(jitter=0.1, dims=(3, 2, 4), hi=(1.5, 0.7, 1.1)):
    m = M.structured_hex(*dims, hi=hi, jitter=jitter, seed=7)
    return m, float(np.prod(hi))


def _state(m, u, p=0.0, ugm=None):
    N = m.nnode
    z = np.zeros((N, 3))
    return dict(dg=z, ug=np.asarray(u, float) + z, ugm=z if ugm is None else ugm + z,
                acg=z, acgm=z, pg=np.full(N, p))


def test_hex_face_blocks_follow_divergence_theorem(oracle_mod):
    O = oracle_mod
    m, vol = _box()
    col, row = O.sparsity(m.nnode, m.ien)
    faces = M.boundary_faces(m)
    assert len(faces[0]) == 2 * (3 * 2 + 3 * 4 + 2 * 4)
    R = O.face_ns(m, np.zeros((m.nnode, 4), np.int32), _state(m, 0.0), PRM, col, row, faces,
                  tensor_flag=4)
    rows = np.repeat(np.arange(m.nnode), np.diff(col))
    for d in range(3):
        gx = np.sum(R["g"][:, d] * m.xg[row, d])
        assert abs(gx - FACT2 * vol) <= 1e-12 * FACT2 * vol, (d, gx, FACT2 * vol)
        # constant test function: the closed surface integral of n vanishes
        assert abs(R["g"][:, d].sum()) <= 1e-13 * FACT2 * vol
    assert np.array_equal(R["d1"], -R["g"])
    assert rows.size == row.size


def test_hex_face_flux_is_outward(oracle_mod):
    O = oracle_mod
    m, vol = _box()
    col, row = O.sparsity(m.nnode, m.ien)
    faces = M.boundary_faces(m)
    ibc = np.zeros((m.nnode, 4), np.int32)
    R = O.face_ns(m, ibc, _state(m, m.xg), PRM, col, row, faces, tensor_flag=2)
    assert abs(R["rhsp"].sum() - 3.0 * vol) <= 1e-12 * vol
    R = O.face_ns(m, ibc, _state(m, [0.3, -1.2, 0.7]), PRM, col, row, faces, tensor_flag=2)
    assert abs(R["rhsp"].sum()) <= 1e-13


def test_hex_face_residual_vanishes_at_the_wall_velocity(oracle_mod):
    O = oracle_mod
    m, _ = _box()
    col, row = O.sparsity(m.nnode, m.ien)
    faces = M.boundary_faces(m)
    wall = np.array([0.2, -0.1, 0.4])
    R = O.face_ns(m, np.zeros((m.nnode, 4), np.int32), _state(m, wall, ugm=wall), PRM, col,
                  row, faces, tensor_flag=2)
    assert np.abs(R["rhsu"]).max() <= 1e-15 and np.abs(R["rhsp"]).max() <= 1e-15


def test_box_faces_are_the_boundary_faces():
    m = M.structured_hex(4, 3, 2, jitter=0.1, seed=3)
    b = M.boundary_masks(m, (0, 0, 0), (1, 1, 1))
    for side in b:
        got = M.box_faces(m, (side,))
        want = M.boundary_faces(m, node_mask=b[side])
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), side
    allf = M.box_faces(m, tuple(b))
    want = M.boundary_faces(m)
    assert np.array_equal(allf[0], want[0]) and np.array_equal(allf[1], want[1])
