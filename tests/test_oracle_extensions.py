"""The oracle's extensions of the reference's element routines, pinned by
closed forms (CPU):

* a constant per-node viscosity scale c gives the reference's assembly with
  mu replaced by c mu (the scale enters as mu sum_a N_a s_a and the shape
  functions sum to one), in both residual forms;
* a constant per-node conductivity scale c gives the temperature assembly
  with kappa_a, kappa_w and kappa_s replaced by c times themselves;
* the scales are local: a scale that differs from 1 only on the nodes of
  elements away from a given row leaves that row unchanged.

The fine-scale residual form itself is pinned bit-exactly against the
reference (tests/test_oracle_golden.py, case ns_tet_vms).
"""
import numpy as np
import pytest

from conftest import golden
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import _lib as L
from oversetdevfort_amd.mesh import Mesh


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("vms", [0, 1])
def test_constant_viscosity_scale_is_scaled_mu(oracle_mod, vms):
    O = oracle_mod
    z = golden("ns_tet_b")
    m = Mesh(xg=z["xg"], ien=z["ien"], elm_id=z["elm_id"])
    col, row = O.sparsity(m.nnode, m.ien)
    st = {k[3:]: z[k] for k in z.files if k.startswith("st_")}
    c = 1.7
    prm = np.array(z["prm"], dtype=np.float64)
    prm_c = prm.copy()
    prm_c[L.PARAMS.index("mu")] *= c
    A = O.assemble_ns(m, z["ibc"], st, prm, col, row, 6, 3, mus=np.full(m.nnode, c), vms=vms)
    B = O.assemble_ns(m, z["ibc"], st, prm_c, col, row, 6, 3, vms=vms)
    U = O.assemble_ns(m, z["ibc"], st, prm, col, row, 6, 3, vms=vms)
    for k in ("k11", "g", "d1", "m", "rhsu", "rhsp"):
        assert _rel(A[k], B[k]) <= 1e-14, k
    assert _rel(A["k11"], U["k11"]) > 1e-6         # the scale is not ignored


def _tem_state(z):
    return {k[3:]: z[k] for k in z.files if k.startswith("st_")}


def test_constant_conductivity_scale_is_scaled_kappa(oracle_mod):
    O = oracle_mod
    z = golden("tem_tet_b")
    m = Mesh(xg=z["xg"], ien=z["ien"], elm_id=z["elm_id"])
    col, row = O.sparsity(m.nnode, m.ien)
    st = _tem_state(z)
    c = 2.5
    q = np.array(z["qprm"], dtype=np.float64)
    qc = q.copy()
    for k in ("kappaa", "kappaw", "kappas"):
        qc[L.QPARAMS.index(k)] *= c
    A = O.assemble_tem(m, z["ibc_t"], st, z["prm"], q, col, row, 6, ks=np.full(m.nnode, c))
    B = O.assemble_tem(m, z["ibc_t"], st, z["prm"], qc, col, row, 6)
    U = O.assemble_tem(m, z["ibc_t"], st, z["prm"], q, col, row, 6)
    for k in ("lhs", "rhs"):
        assert _rel(A[k], B[k]) <= 1e-14, k
    assert _rel(A["lhs"], U["lhs"]) > 1e-12


def test_scales_are_local(oracle_mod):
    """Raise the viscosity scale on one corner node of a hex cavity: against
    a scale of one everywhere, rows of nodes that share no element with it
    are unchanged, bit for bit."""
    O = oracle_mod
    prob = C.cavity(4, seed=3, jitter=0.05)
    m, f = prob.mesh, prob.fields
    N = m.nnode
    col, row = O.sparsity(N, m.ien)
    rng = np.random.default_rng(3)
    st = dict(dg=np.zeros((N, 3)), ug=0.3 * rng.normal(size=(N, 3)), ugm=np.zeros((N, 3)),
              acg=rng.normal(size=(N, 3)), acgm=np.zeros((N, 3)), pg=rng.normal(size=N))
    s = np.ones(N)
    s[0] = 3.0
    near = np.zeros(N, bool)
    near[np.unique(m.ien[np.any(m.ien == 0, axis=1)])] = True
    far = ~near
    prm = O.params_array(prob.params)
    A = O.assemble_ns(m, prob.ibc, st, prm, col, row, 6, 1, mus=s, vms=1)
    B = O.assemble_ns(m, prob.ibc, st, prm, col, row, 6, 1, mus=np.ones(N), vms=1)
    assert np.array_equal(A["rhsu"][far], B["rhsu"][far])
    assert not np.array_equal(A["rhsu"][near], B["rhsu"][near])
    rows = np.repeat(np.arange(N), np.diff(col))
    assert np.array_equal(A["k11"][far[rows]], B["k11"][far[rows]])


def test_colour_ordered_assembly(oracle_mod):
    """The colour-by-colour assembly of the host baseline: with one element
    per colour in index order it is the element-order assembly bit for bit;
    with the 8-colouring of the hex grid only the summation order per entry
    changes, so it agrees to rounding (1e-14 normwise) with the same
    Dirichlet rows, CFL and negative-Jacobian count."""
    from oversetdevfort_amd import mesh as M
    O = oracle_mod
    prob = C.cavity(6, seed=4, jitter=0.05)
    m = prob.mesh
    N = m.nnode
    col, row = O.sparsity(N, m.ien)
    rng = np.random.default_rng(4)
    st = dict(dg=0.01 * rng.normal(size=(N, 3)), ug=0.3 * rng.normal(size=(N, 3)),
              ugm=np.zeros((N, 3)), acg=rng.normal(size=(N, 3)), acgm=np.zeros((N, 3)),
              pg=rng.normal(size=N))
    prm = O.params_array(prob.params)
    ref = O.assemble_ns(m, prob.ibc, st, prm, col, row, 6, 1)
    keys = ("k11", "g", "d1", "m", "rhsu", "rhsp")
    one = O.assemble_ns_col(m, prob.ibc, st, prm, col, row, np.arange(m.nelem), 6, 1)
    for k in keys:
        assert np.array_equal(one[k], ref[k]), k
    assert one["cfl"] == ref["cfl"] and one["nneg"] == ref["nneg"]
    colors = M.hex_colors(m)
    assert len(np.unique(colors)) == 8
    col8 = O.assemble_ns_col(m, prob.ibc, st, prm, col, row, colors, 6, 1)
    for k in keys:
        a, b = col8[k].ravel(), ref[k].ravel()
        assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(b), k
    assert not all(np.array_equal(col8[k], ref[k]) for k in keys)
    assert col8["cfl"] == ref["cfl"] and col8["nneg"] == ref["nneg"]
    vec = O.assemble_ns_col(m, prob.ibc, st, prm, col, row, colors, 2, 1)
    assert np.array_equal(vec["rhsu"], col8["rhsu"])
    assert not vec["k11"].any()
