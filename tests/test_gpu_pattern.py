"""genSparStruc on the device (genSparStruc.f90:4-105) against the oracle,
whose col/row are bit-exact against the reference (tests/test_oracle_golden.py):

* col and row bit for bit on a hex mesh whose nodes and elements are
  numbered at random (rows no longer banded), and on a tet mesh;
* the maps built beside the pattern: the scatter map emap through a full
  assembly of the renumbered mesh against the oracle (1e-12), and the
  transpose map through LHS^T products against the transposed oracle matrix.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-300)


def _renumbered(seed=5):
    rng = np.random.default_rng(seed)
    prob = C.cavity(5, seed=seed, jitter=0.1)
    m = prob.mesh
    N, E = m.nnode, m.nelem
    perm = rng.permutation(N)              # new id of old node n: perm[n]
    inv = np.argsort(perm)
    eperm = rng.permutation(E)
    ien = perm[m.ien[eperm]].astype(np.int32)
    mesh = M.Mesh(xg=m.xg[inv], ien=ien, elm_id=m.elm_id[eperm])
    prob.mesh = mesh
    prob.ibc = prob.ibc[inv]
    prob.bcval = prob.bcval[inv]
    prob.fields = {k: v[inv] for k, v in prob.fields.items()}
    return prob


def test_device_pattern_bit_exact(oracle_mod):
    from oversetdevfort_amd import solver as S
    O = oracle_mod
    prob = _renumbered()
    ctx = S.Context(0)
    ctx.upload_mesh(prob.mesh)
    col, row = ctx.genSparStruc()
    oc, orow = O.sparsity(prob.mesh.nnode, prob.mesh.ien)
    assert np.array_equal(col, oc) and np.array_equal(row, orow)
    ctx.close()
    tm = M.hex_to_tets(M.structured_hex(4, 3, 3, jitter=0.1, seed=2))
    ctx = S.Context(0)
    ctx.upload_mesh(tm)
    col, row = ctx.genSparStruc()
    oc, orow = O.sparsity(tm.nnode, tm.ien)
    assert np.array_equal(col, oc) and np.array_equal(row, orow)
    ctx.close()


def test_renumbered_assembly_and_transpose(oracle_mod):
    from oversetdevfort_amd import solver as S
    O = oracle_mod
    prob = _renumbered(seed=8)
    m = prob.mesh
    N = m.nnode
    rng = np.random.default_rng(3)
    st = dict(dg=np.zeros((N, 3)), ug=rng.normal(size=(N, 3)), ugm=np.zeros((N, 3)),
              acg=rng.normal(size=(N, 3)), acgm=np.zeros((N, 3)), pg=rng.normal(size=N))
    ctx = S.setup(prob)
    for k in ("dg", "ug", "ugm", "acg", "acgm", "pg"):
        ctx.set_field(k.upper(), st[k])
        ctx.set_field(k.upper() + "OLD", st[k])
    ctx.assembleNavSto(6, 1)
    col, row = O.sparsity(N, m.ien)
    ref = O.assemble_ns(m, prob.ibc, st, O.params_array(prob.params), col, row, 6, 1)
    A = ctx.get_lhs()
    for k in ("k11", "g", "d1", "m"):
        assert _rel(A[k], ref[k]) <= 1e-12, k
    ru, rp = ctx.get_rhs()
    assert _rel(ru, ref["rhsu"]) <= 1e-12 and _rel(rp, ref["rhsp"]) <= 1e-12
    # LHS^T through the transpose map
    B = np.zeros((row.size, 4, 4))
    B[:, :3, :3] = ref["k11"].reshape(-1, 3, 3)
    B[:, :3, 3] = ref["g"]
    B[:, 3, :3] = ref["d1"]
    B[:, 3, 3] = ref["m"]
    At = sp.bsr_matrix((B, row, col), shape=(4 * N, 4 * N)).T.tocsr()
    xu, xp = rng.normal(size=(N, 3)), rng.normal(size=N)
    yu, yp = ctx.SparseProdUP_3D_T(xu, xp)
    y = At @ np.concatenate([xu, xp[:, None]], axis=1).ravel()
    y = y.reshape(N, 4)
    assert _rel(yu, y[:, :3]) <= 1e-13 and _rel(yp, y[:, 3]) <= 1e-13
    ctx.close()
