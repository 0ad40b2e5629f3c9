"""The quenching solver's element kernels on linear tets -- the element of
the reference's live dispatch -- on the GPU against the reference's own
output: the fixtures hold IntElmAss_NSVOF_Quenching_STAB with
BCLhs_NS_conv + FillSparseMat_NS_conv and IntElmAss_Tem_Quenching_STAB
with BCLhs_tem + FillSparseMat_tem as the compiled reference computed them
(relinked harness), on tet meshes of fluid and solid elements; case b turns
discontinuity capturing on.  Every LHS array and RHS to 1e-12 (normwise),
the products to 1e-13, the pattern bit for bit."""
import numpy as np
import pytest

from conftest import golden
from oversetdevfort_amd import _lib as L
from oversetdevfort_amd.mesh import Mesh
from oversetdevfort_amd.solver import NSVOF_BLOCKS

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-300)


def _ctx(z, ibc8, lsc_kdc=0.0):
    """A context whose alpha stage is the fixture's state: every field and
    its old copy are set to it."""
    from oversetdevfort_amd import solver as S
    m = Mesh(xg=z["xg"], ien=z["ien"], elm_id=z["elm_id"])
    ctx = S.Context(0)
    ctx.upload_mesh(m)
    col, row = ctx.genSparStruc()
    assert np.array_equal(col, z["col"]) and np.array_equal(row, z["row"])
    almi, alfi, gami, delt = (float(v) for v in z["prm"])
    grav = tuple(float(v) for v in z["ex"][:3]) if "ex" in z.files else (0.0, 0.0, -9.81)
    ctx.set_params(dict(almi=almi, alfi=alfi, gami=gami, beti=0.25, delt=delt, rho=1.0,
                        mu=1.0, kdc=0.0, grav=grav))
    q = dict(zip(L.QPARAMS[:19], (float(v) for v in z["qprm"])))
    q.update(fluid_id=102, lsc_kdc=lsc_kdc)
    ctx.set_quench_params(q)
    ctx.set_ibc(ibc8)
    for k in z.files:
        if k.startswith("st_"):
            f = k[3:].upper()
            ctx.set_field(f, z[k])
            ctx.set_field(f + "OLD", z[k])
    return ctx


@pytest.mark.parametrize("name", ["nsvof_tet_a", "nsvof_tet_b"])
def test_nsvof_tet_matches_reference(name):
    z = golden(name)
    print(f"NSVOF tets {name}: {z['ien'].shape[0]} elements", flush=True)
    ctx = _ctx(z, z["ibc"], lsc_kdc=float(z["ex"][3]))
    ctx.assemble_nsvof(int(z["tensor_flag"]))
    A, rhs = ctx.get_nsvof_system()
    for k in NSVOF_BLOCKS:
        assert _rel(A[k], z[k]) <= 1e-12, (k, _rel(A[k], z[k]))
    for k in ("rhsu", "rhsp", "rhsls"):
        assert _rel(rhs[k], z[k]) <= 1e-12, (k, _rel(rhs[k], z[k]))
    yu, yp, yl = ctx.SparseProd_NS_conv(z["xu"], z["xp"], z["xl"], L.OFD_FP64)
    for a, b in ((yu, z["yu"]), (yp, z["yp"]), (yl, z["yl"])):
        assert _rel(a, b) <= 1e-13
    ctx.close()


@pytest.mark.parametrize("name", ["tem_tet_a", "tem_tet_b"])
def test_tem_tet_matches_reference(name):
    z = golden(name)
    print(f"temperature tets {name}: {z['ien'].shape[0]} elements", flush=True)
    N = z["xg"].shape[0]
    ibc8 = np.zeros((N, 8), np.int32)
    ibc8[:, 5] = z["ibc_t"]
    ctx = _ctx(z, ibc8)
    ctx.assemble_tem(int(z["tensor_flag"]))
    lhs, rhs = ctx.get_tem_system()
    assert _rel(lhs, z["lhs"]) <= 1e-12, _rel(lhs, z["lhs"])
    assert _rel(rhs, z["rhs"]) <= 1e-12, _rel(rhs, z["rhs"])
    assert _rel(ctx.SparseProd_LS(z["x"], L.OFD_FP64), z["y"]) <= 1e-13
    ctx.close()
