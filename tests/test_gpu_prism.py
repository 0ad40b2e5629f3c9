"""The NS assembly on linear prisms (linprishl and the 6-point rule of
genGPandGWtetpri, eval_shape_FEM.f90:438-477, genGP_GW.f90:317-430), the
third element of the reference's live dispatch (eval_shape_FEM.f90:18-26),
on the GPU against the oracle, whose prism shape functions and rule are
bit-exact against the compiled reference: jittered hexes split into
prisms, random state and Dirichlet flags, two element ids; VEC+MAT, VEC and
MAT, one or both subdomains, to 1e-12 (normwise) and 1e-11 per entry; the
pattern bit for bit; the block product to 1e-13."""
import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import mesh as M

pytestmark = pytest.mark.gpu

PRM = dict(almi=0.8333333333333334, alfi=0.6666666666666666, gami=0.6666666666666667,
           beti=0.34722222222222227, delt=0.01, rho=1.3, mu=0.02, kdc=0.0,
           grav=(0.1, -0.2, -9.81))


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("flag,sub,seed", [(6, 1, 1), (6, 3, 2), (2, 1, 3), (4, 2, 4)])
def test_prism_assembly_matches_oracle(oracle_mod, flag, sub, seed):
    from oversetdevfort_amd import solver as S
    O = oracle_mod
    rng = np.random.default_rng(seed)
    m = M.hex_to_prisms(M.structured_hex(4, 3, 3, hi=(1.0, 0.9, 1.1), jitter=0.08, seed=seed))
    m.elm_id = rng.integers(1, 3, size=m.nelem).astype(np.int32)
    N = m.nnode
    ibc = (rng.uniform(size=(N, 4)) < 0.15).astype(np.int32)
    st = {k: rng.normal(size=(N, 3)) for k in ("ug", "ugm", "acg", "acgm")}
    st["dg"] = 1e-3 * rng.normal(size=(N, 3))
    st["pg"] = rng.normal(size=N)
    print(f"prisms flag {flag} sub {sub} seed {seed}: {m.nelem} prisms, {N} nodes", flush=True)
    ctx = S.Context(0)
    ctx.upload_mesh(m)
    col, row = ctx.genSparStruc()
    ocol, orow = O.sparsity(N, m.ien)
    assert np.array_equal(col, ocol) and np.array_equal(row, orow)
    ctx.set_params(PRM)
    ctx.set_ibc(ibc)
    for k, v in st.items():
        ctx.set_field(k.upper(), v)
        ctx.set_field(k.upper() + "OLD", v)
    cfl = ctx.assembleNavSto(flag, sub)
    R = O.assemble_ns(m, ibc, st, O.params_array(PRM), ocol, orow, flag, sub)
    assert R["nneg"] == 0
    if flag & 4:
        A = ctx.get_lhs()
        for k in ("k11", "g", "d1", "m"):
            assert _rel(A[k], R[k]) <= 1e-12, (k, _rel(A[k], R[k]))
            assert np.abs(A[k] - R[k]).max() <= 1e-11 * max(np.abs(R[k]).max(), 1e-300), k
        xu, xp = rng.normal(size=(N, 3)), rng.normal(size=N)
        yu, yp = ctx.SparseProdUP_3D(xu, xp, L.OFD_FP64)
        ou, op = O.spmv_up(ocol, orow, A, xu, xp)
        assert _rel(yu, ou) <= 1e-13 and _rel(yp, op) <= 1e-13
    if flag & 2:
        ru, rp = ctx.get_rhs()
        assert _rel(ru, R["rhsu"]) <= 1e-12 and _rel(rp, R["rhsp"]) <= 1e-12
    assert abs(cfl - R["cfl"]) <= 1e-12 * R["cfl"]
    ctx.close()
