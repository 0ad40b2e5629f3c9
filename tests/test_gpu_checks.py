"""Argument and state checks at the C ABI, and the NS assembly on an
unstructured tet mesh.

* An unstructured (Delaunay) tet mesh whose first-fit element colouring
  needs more than 32 colours: the pattern bit for bit against the oracle's
  genSparStruc restatement and the NS assembly (IntElmAss_NS_overset +
  BCLhs_3D + FillSparseMat_3D) to 1e-12 against the oracle.
* A mesh upload drops the weak faces registered for the previous mesh; tet
  faces take FACE_OR 1..4, prism faces are refused.
* A right-hand side with a NaN is reported as OFD_ERR_NONFINITE by every
  solver entry (SparseGMRES_up, SparseGMRES_tem, SparseGMRES_ns_conv, the
  restarted solve and the adjoint's flexible GMRES), never as a zero
  solution.
* ofd_check_pattern, which the Fortran drop-ins call with their col/row
  arguments, accepts the context's pattern and refuses any other.
"""
import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd.mesh import Mesh, structured_hex

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-300)


def _delaunay_tets(n, seed):
    """Delaunay tets of a jittered n^3 grid of points in the unit cube,
    positively oriented for lintetshl (det[x1-x4, x2-x4, x3-x4] > 0), flat
    boundary slivers dropped."""
    import scipy.spatial as sp
    m = structured_hex(n, n, n, jitter=0.3, seed=seed)
    x = m.xg
    ien = sp.Delaunay(x).simplices.astype(np.int32)
    det = np.linalg.det(np.stack([x[ien[:, 0]] - x[ien[:, 3]], x[ien[:, 1]] - x[ien[:, 3]],
                                  x[ien[:, 2]] - x[ien[:, 3]]], axis=2))
    keep = np.abs(det) > 1e-3 / n ** 3
    ien, det = ien[keep], det[keep]
    neg = det < 0
    ien[neg, 0], ien[neg, 1] = ien[neg, 1].copy(), ien[neg, 0].copy()
    used = np.unique(ien)
    remap = np.full(x.shape[0], -1, np.int64)
    remap[used] = np.arange(used.size)
    return Mesh(xg=x[used].copy(), ien=remap[ien].astype(np.int32),
                elm_id=np.ones(ien.shape[0], np.int32))


def _colours(m):
    masks = [set() for _ in range(m.nnode)]
    nc = 0
    for nodes in m.ien:
        used = set().union(*[masks[n] for n in nodes])
        c = 0
        while c in used:
            c += 1
        for n in nodes:
            masks[n].add(c)
        nc = max(nc, c + 1)
    return nc


def test_unstructured_tet_mesh_beyond_32_colours(oracle_mod):
    from oversetdevfort_amd import solver as S
    O = oracle_mod
    m = _delaunay_tets(12, seed=5)
    ncol = _colours(m)
    print(f"delaunay tets: {m.nelem} tets, {m.nnode} nodes, {ncol} colours", flush=True)
    assert ncol > 32
    N = m.nnode
    rng = np.random.default_rng(7)
    ctx = S.Context(0)
    ctx.upload_mesh(m)
    col, row = ctx.genSparStruc()
    ocol, orow = O.sparsity(N, m.ien)
    assert np.array_equal(col, ocol) and np.array_equal(row, orow)
    p = C.cavity(2).params
    ctx.set_params(p)
    ibc = np.zeros((N, 4), np.int32)
    ibc[rng.uniform(size=N) < 0.1, :3] = 1
    ibc[rng.uniform(size=N) < 0.05, 3] = 1
    ctx.set_ibc(ibc)
    st = {k: rng.normal(size=(N, 3)) for k in ("ug", "ugm", "acg", "acgm")}
    st["dg"] = np.zeros((N, 3))    # the tets keep the orientation they were given
    st["pg"] = rng.normal(size=N)
    for k, v in st.items():
        ctx.set_field(k.upper(), v)
        ctx.set_field(k.upper() + "OLD", v)   # the alpha stage is the state
    ctx.assembleNavSto(6, 1)
    A = ctx.get_lhs()
    ru, rp = ctx.get_rhs()
    R = O.assemble_ns(m, ibc, st, O.params_array(p), ocol, orow, 6, 1)
    assert _rel(ru, R["rhsu"]) <= 1e-12 and _rel(rp, R["rhsp"]) <= 1e-12
    for k in ("k11", "g", "d1", "m"):
        assert _rel(A[k], R[k]) <= 1e-12, (k, _rel(A[k], R[k]))
    ctx.close()


def test_upload_drops_weak_faces():
    from oversetdevfort_amd import solver as S
    from oversetdevfort_amd.mesh import box_faces, hex_to_prisms, hex_to_tets
    hexm = structured_hex(3, 3, 3, jitter=0.05, seed=2)
    ctx = S.Context(0)
    ctx.upload_mesh(hexm)
    ctx.build_pattern()
    ctx.set_params(C.cavity(2).params)
    ctx.set_weak_faces(*box_faces(hexm, ("zhi",)))
    ctx.assembleNavSto(6, 1)
    tets = hex_to_tets(hexm)
    ctx.upload_mesh(tets)          # the faces named hexes of the old mesh
    ctx.build_pattern()
    ctx.assembleNavSto(6, 1)       # no faces left, so no face launch
    # tets take FACE_OR 1..4 (the reference's face map); prisms have no faces
    with pytest.raises(L.OfdError) as e:
        ctx.set_weak_faces(np.array([0], np.int32), np.array([5], np.int32))
    assert e.value.code == L.OFD_ERR_ARG
    ctx.set_weak_faces(np.array([0], np.int32), np.array([1], np.int32))
    ctx.upload_mesh(hex_to_prisms(hexm))
    ctx.build_pattern()
    with pytest.raises(L.OfdError) as e:
        ctx.set_weak_faces(np.array([0], np.int32), np.array([1], np.int32))
    assert e.value.code == L.OFD_ERR_UNSUPPORTED
    ctx.close()


def test_nonfinite_right_hand_sides_are_errors():
    from oversetdevfort_amd import solver as S
    prob = C.cavity(3, seed=1)
    N = prob.mesh.nnode
    ctx = S.setup(prob)
    ctx.assembleNavSto(6, 1)
    ru, rp = ctx.get_rhs()
    rp = rp.copy()
    rp[N // 2] = np.nan
    ctx.set_rhs(ru, rp)
    for call in (lambda: ctx.SparseGMRES_up(1e-6, 30),
                 lambda: ctx.solve(rtol=1e-8, max_cycles=5)):
        with pytest.raises(L.OfdError) as e:
            call()
        assert e.value.code == L.OFD_ERR_NONFINITE
    gu = np.zeros((N, 3))
    gu[0, 0] = np.nan
    with pytest.raises(L.OfdError) as e:
        ctx.solve_adjoint(gu, np.zeros(N), max_cycles=2)
    assert e.value.code == L.OFD_ERR_NONFINITE
    ctx.close()
    prob = C.quench(3, seed=3, jitter=0.05)
    ctx = S.setup(prob)
    ctx.set_quench_params(prob.material)
    ctx.assemble_tem(6)
    lhs, rhs = ctx.get_tem_system()
    rhs = rhs.copy()
    rhs[1] = np.nan
    ctx.set_tem_system(rhs=rhs)
    with pytest.raises(L.OfdError) as e:
        ctx.SparseGMRES_tem(1e-6, 30)
    assert e.value.code == L.OFD_ERR_NONFINITE
    ctx.assemble_nsvof(6)
    A, r = ctx.get_nsvof_system()
    r["rhsls"] = r["rhsls"].copy()
    r["rhsls"][2] = np.inf
    ctx.set_nsvof_system(rhs=r)
    with pytest.raises(L.OfdError) as e:
        ctx.SparseGMRES_ns_conv(1e-6, 30)
    assert e.value.code == L.OFD_ERR_NONFINITE
    ctx.close()


def test_check_pattern_refuses_other_patterns():
    from oversetdevfort_amd import solver as S
    prob = C.cavity(3, seed=1)
    ctx = S.setup(prob)
    col, row = ctx.genSparStruc()
    ctx.check_pattern(col, row)
    for c2, r2 in ((col, np.roll(row, 1)), (col[:-1], row), (col, row[:-1])):
        with pytest.raises(L.OfdError) as e:
            ctx.check_pattern(c2, r2)
        assert e.value.code == L.OFD_ERR_ARG
    r3 = row.copy()
    r3[len(r3) // 2] += 1
    with pytest.raises(L.OfdError) as e:
        ctx.check_pattern(col, r3)
    assert e.value.code == L.OFD_ERR_ARG
    ctx.close()
