"""Weak-Dirichlet boundary faces on the GPU against the oracle.

assembleNavSto adds FaceAssembly_NS_weak and FaceAssembly_NS_outflow over
the BCugType-2 faces after the volume loop (assembleNavSto.f90:50-60); so
does ofd_assemble_ns for the faces given to ofd_set_weak_faces.  The oracle's
face routine is bit-exact against the reference on tet faces
(tests/test_oracle_faces.py); on hex faces (the face map is an extension) it
is pinned by closed forms.  Here:

* the assembled LHS and RHS (volume + faces) of a jittered 6^3 cavity with
  every boundary face weak, a random state with a moving mesh velocity,
  random constraints (whose diagonal collects 2 per face: each face passes
  through BCLhs_3D once per routine) and two element ids, for flags 6/2/4
  and subdomain flags 1/2/3 (the outflow routine ignores the subdomain) --
  to 1e-12 normwise;
* a time step of the weak-wall cavity (lid strong, five walls weak), two
  Newton iterations, against the oracle's assembly and a direct solve --
  fields to 1e-10;
* linear tets (the reference's own face map, 3-point faces): the same
  comparison on a jittered, Kuhn-split 3^3 box with every boundary face weak,
  and the reference's own face output (fixtures face_tet_a, face_tet_vec)
  added to the oracle's volume assembly, to 1e-12;
* bad face lists are refused.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def S():
    from oversetdevfort_amd import solver
    return solver


def _random_case(seed=11, n=6):
    rng = np.random.default_rng(seed)
    prob = C.cavity(n, seed=seed, jitter=0.1)
    m = prob.mesh
    N = m.nnode
    m.elm_id[:] = rng.integers(1, 3, size=m.nelem)
    ibc = (rng.uniform(size=(N, 4)) < 0.1).astype(np.int32)
    st = dict(dg=0.01 * rng.normal(size=(N, 3)), ug=rng.normal(size=(N, 3)),
              ugm=0.3 * rng.normal(size=(N, 3)), acg=rng.normal(size=(N, 3)),
              acgm=np.zeros((N, 3)), pg=rng.normal(size=N))
    return prob, ibc, st, M.boundary_faces(m)


@pytest.mark.parametrize("tflag,sflag", [(6, 1), (6, 3), (2, 2), (4, 1), (6, 2)])
def test_face_assembly_matches_oracle(S, oracle_mod, tflag, sflag):
    O = oracle_mod
    prob, ibc, st, faces = _random_case()
    m = prob.mesh
    col, row = O.sparsity(m.nnode, m.ien)
    prm = O.params_array(prob.params)
    ctx = S.setup(prob)
    ctx.set_ibc(ibc)
    for k in ("dg", "ug", "ugm", "acg", "acgm", "pg"):
        ctx.set_field(k.upper(), st[k])
        ctx.set_field(k.upper() + "OLD", st[k])      # alpha stage = the state
    ctx.set_weak_faces(*faces)
    ctx.assembleNavSto(tflag, sflag)
    A, (ru, rp) = ctx.get_lhs(), ctx.get_rhs()
    ref = O.assemble_ns(m, ibc, st, prm, col, row, tflag, sflag)
    ref = O.face_ns(m, ibc, st, prm, col, row, faces, tflag, sflag, out=ref)
    bare = O.assemble_ns(m, ibc, st, prm, col, row, tflag, sflag)
    if tflag & 4:
        for k in ("k11", "g", "d1", "m"):
            assert _rel(A[k], ref[k]) <= 1e-12, (k, _rel(A[k], ref[k]))
        assert _rel(bare["g"], ref["g"]) > 1e-6          # the faces add to G
    if tflag & 2:
        assert _rel(ru, ref["rhsu"]) <= 1e-12 and _rel(rp, ref["rhsp"]) <= 1e-12
        assert _rel(bare["rhsu"], ref["rhsu"]) > 1e-6
    # clearing the faces gives the volume assembly back
    ctx.set_weak_faces([], [])
    ctx.assembleNavSto(tflag, sflag)
    if tflag & 2:
        assert _rel(ctx.get_rhs()[0], bare["rhsu"]) <= 1e-12
    ctx.close()


def test_weak_cavity_time_step(S, oracle_mod):
    O = oracle_mod
    prob = C.weak_cavity(6, seed=2, jitter=0.05)
    m, p = prob.mesh, prob.params
    N = m.nnode
    col, row = O.sparsity(N, m.ien)
    prm = O.params_array(p)
    ga, dt, af = p["gami"], p["delt"], p["alfi"]
    st = {k: v.copy() for k, v in prob.fields.items()}
    ctx = S.setup(prob)
    ctx.predictor()
    ctx.apply_dirichlet()
    st["ug"] = st["ugold"].copy()
    st["acg"] = (ga - 1.0) / ga * st["acgold"]
    st["pg"] = st["pgold"].copy()
    st["ug"], st["acg"] = C.apply_dirichlet(st["ug"], st["acg"], st["ugold"], st["acgold"],
                                            prob)
    for it in range(2):
        a = {k: st[k + "old"] + af * (st[k] - st[k + "old"]) for k in ("dg", "ug", "ugm")}
        a.update(acg=st["acgold"] + p["almi"] * (st["acg"] - st["acgold"]),
                 acgm=st["acgmold"] + p["almi"] * (st["acgm"] - st["acgmold"]), pg=st["pg"])
        R = O.assemble_ns(m, prob.ibc, a, prm, col, row, 6, 1)
        R = O.face_ns(m, prob.ibc, a, prm, col, row, prob.weak_faces, 6, 1, out=R)
        B = np.zeros((row.size, 4, 4))
        B[:, :3, :3] = R["k11"].reshape(-1, 3, 3)
        B[:, :3, 3] = R["g"]
        B[:, 3, :3] = R["d1"]
        B[:, 3, 3] = R["m"]
        Asp = sp.bsr_matrix((B, row, col), shape=(4 * N, 4 * N)).tocsc()
        b = np.concatenate([R["rhsu"], R["rhsp"][:, None]], axis=1).ravel()
        x = spla.spsolve(Asp, b).reshape(N, 4)
        st["acg"] = st["acg"] + x[:, :3]
        st["ug"] = st["ug"] + ga * dt * x[:, :3]
        st["pg"] = st["pg"] + af * ga * dt * x[:, 3]
        info = S.newton_step(ctx, rtol=1e-13, precision=L.OFD_FP64, kspace=30,
                             max_cycles=800)
        assert info["converged"], info["relres"]
    errs = {k: _rel(ctx.get_field(k.upper()), st[k]) for k in ("ug", "acg", "pg")}
    print(f"weak-wall cavity 6^3, one step vs oracle: " +
          ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-10, errs
    # the weak walls hold the velocity near zero without IBC
    wall = M.boundary_masks(m, (0, 0, 0), (1, 1, 1))["zlo"]
    assert np.abs(ctx.get_field("UG")[wall]).max() < 0.1 * np.abs(st["ug"]).max()
    ctx.close()


def _tet_case(seed=13):
    rng = np.random.default_rng(seed)
    hexm = M.structured_hex(3, 3, 3, jitter=0.08, seed=seed)
    m = M.hex_to_tets(hexm)
    N = m.nnode
    m.elm_id[:] = rng.integers(1, 3, size=m.nelem)
    ibc = (rng.uniform(size=(N, 4)) < 0.1).astype(np.int32)
    st = dict(dg=0.005 * rng.normal(size=(N, 3)), ug=rng.normal(size=(N, 3)),
              ugm=0.3 * rng.normal(size=(N, 3)), acg=rng.normal(size=(N, 3)),
              acgm=np.zeros((N, 3)), pg=rng.normal(size=N))
    return m, ibc, st, M.boundary_faces(m)


def _tet_ctx(S, m, ibc, st, prm_dict):
    ctx = S.Context(0)
    ctx.upload_mesh(m)
    ctx.build_pattern()
    ctx.set_params(prm_dict)
    ctx.set_ibc(ibc)
    for k in ("dg", "ug", "ugm", "acg", "acgm", "pg"):
        ctx.set_field(k.upper(), st[k])
        ctx.set_field(k.upper() + "OLD", st[k])      # alpha stage = the state
    return ctx


@pytest.mark.parametrize("tflag,sflag", [(6, 1), (6, 3), (2, 2), (4, 3)])
def test_tet_face_assembly_matches_oracle(S, oracle_mod, tflag, sflag):
    O = oracle_mod
    m, ibc, st, faces = _tet_case()
    p = C.cavity(2).params
    col, row = O.sparsity(m.nnode, m.ien)
    prm = O.params_array(p)
    ctx = _tet_ctx(S, m, ibc, st, p)
    ctx.set_weak_faces(*faces)
    ctx.assembleNavSto(tflag, sflag)
    A, (ru, rp) = ctx.get_lhs(), ctx.get_rhs()
    ref = O.assemble_ns(m, ibc, st, prm, col, row, tflag, sflag)
    ref = O.face_ns(m, ibc, st, prm, col, row, faces, tflag, sflag, ngb=3, out=ref)
    bare = O.assemble_ns(m, ibc, st, prm, col, row, tflag, sflag)
    errs = {}
    if tflag & 4:
        for k in ("k11", "g", "d1", "m"):
            errs[k] = _rel(A[k], ref[k])
        assert _rel(bare["g"], ref["g"]) > 1e-6          # the faces add to G
    if tflag & 2:
        errs["rhsu"], errs["rhsp"] = _rel(ru, ref["rhsu"]), _rel(rp, ref["rhsp"])
        assert _rel(bare["rhsu"], ref["rhsu"]) > 1e-6
    print(f"tet faces ({len(faces[0])}), flags {tflag}/{sflag}: " +
          ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-12, errs
    ctx.close()


@pytest.mark.parametrize("name", ["face_tet_a", "face_tet_vec"])
def test_tet_faces_replay_reference_fixtures(S, oracle_mod, name):
    """The reference's own face routines' output (3-point faces; the 1-point
    fixture is a face rule the reference's input never selects, NGAUSSB =
    NSHLB = 3) added to the oracle's volume assembly equals the GPU's volume +
    face assembly."""
    from conftest import golden
    O = oracle_mod
    z = golden(name)
    m = M.Mesh(xg=z["xg"], ien=z["ien"], elm_id=z["elm_id"])
    st = {k[3:]: z[k] for k in z.files if k.startswith("st_")}
    prm = z["prm"]
    p = dict(almi=prm[0], alfi=prm[1], gami=prm[2], beti=prm[3], delt=prm[4], rho=prm[5],
             mu=prm[6], kdc=prm[7], grav=tuple(prm[8:11]))
    tflag, sflag = int(z["tensor_flag"]), int(z["subdomain_flag"])
    col, row = O.sparsity(m.nnode, m.ien)
    ctx = _tet_ctx(S, m, z["ibc"], st, p)
    ctx.set_subdomains(1, 2)
    ctx.set_weak_faces(z["f2e"], z["face_or"])
    ctx.assembleNavSto(tflag, sflag)
    vol = O.assemble_ns(m, z["ibc"], st, prm, col, row, tflag, sflag)
    errs = {}
    if tflag & 4:
        A = ctx.get_lhs()
        for k in ("k11", "g", "d1", "m"):
            errs[k] = _rel(A[k], vol[k] + z[k])
    ru, rp = ctx.get_rhs()
    errs["rhsu"] = _rel(ru, vol["rhsu"] + z["rhsu"])
    errs["rhsp"] = _rel(rp, vol["rhsp"] + z["rhsp"])
    print(f"{name}: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-12, errs
    ctx.close()


def test_bad_face_lists_are_refused(S):
    prob = C.cavity(3)
    ctx = S.setup(prob)
    E = prob.mesh.nelem
    for f2e, fo in (([E], [1]), ([0], [7]), ([0, 0], [2, 2]), ([-1], [1])):
        with pytest.raises(L.OfdError):
            ctx.set_weak_faces(f2e, fo)
    ctx.close()
