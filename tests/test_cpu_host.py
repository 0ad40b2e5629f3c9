"""CPU-only checks: the C-ABI library loads and exports every symbol of
include/ofd.h (no compute call without a GPU), the oracle's hex-path pieces
against closed forms the mathematics fixes, and the host-side set-up."""
import ctypes
import os

import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M


def test_library_exports_every_header_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libofd.so not built (run __graft_entry__.build())")
    names = L.header_symbols()
    assert len(names) >= 25
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_product_path_refuses_missing_library(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        L._LIB_SAVE = L._LIB
        try:
            L._LIB = None
            L.load(str(tmp_path / "libofd.so"))
        finally:
            L._LIB = L._LIB_SAVE


def test_hex_shape_partition_of_unity_and_kronecker(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(0)
    for pt in rng.uniform(-1, 1, size=(20, 3)):
        shl, sg = O.shape_ref(8, pt)
        assert abs(shl.sum() - 1.0) < 1e-15
        assert np.abs(sg.sum(axis=0)).max() < 1e-15
    # N_a(x_b) = delta_ab at the corners of linhexshl's node order
    for b, c in enumerate(M.HEX_CORNERS):
        shl, _ = O.shape_ref(8, 2.0 * c - 1.0)
        e = np.zeros(8)
        e[b] = 1.0
        assert np.array_equal(shl, e)


def test_hex_rule_integrates_cubics(oracle_mod):
    """The 2x2x2 rule is exact for x^i y^j z^k, i,j,k <= 3; the truncated
    abscissa of the reference leaves an O(1e-14) error, no more."""
    gp, gw = oracle_mod.gauss_volume(8)
    for i in range(4):
        for j in range(4):
            for k in range(4):
                q = np.sum(gw * gp[:, 0] ** i * gp[:, 1] ** j * gp[:, 2] ** k)
                ex = np.prod([0.0 if n % 2 else 2.0 / (n + 1) for n in (i, j, k)])
                assert abs(q - ex) < 5e-14


def test_hex_volume_identity(oracle_mod):
    """sum_q DetJ w = |element| for a trilinear hex (exact for the 2-pt rule)."""
    O = oracle_mod
    m = M.structured_hex(2, 2, 2, jitter=0.08, seed=5)
    gp, gw = O.gauss_volume(8)
    total = 0.0
    for e in range(m.nelem):
        xl = m.xg[m.ien[e]]
        vol = 0.0
        for q in range(8):
            *_, dj = O.eval_shape_fem(8, gp[q], xl, np.zeros_like(xl))
            assert dj > 0
            vol += dj * gw[q]
        total += vol
    assert abs(total - 1.0) < 1e-12


def test_hex_gradients_exact_for_linear_field(oracle_mod):
    O = oracle_mod
    m = M.structured_hex(1, 1, 1, jitter=0.0)
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3))
    xl = m.xg[m.ien[0]] @ (np.eye(3) + 0.2 * A).T  # affine element
    f = xl @ np.array([1.5, -2.0, 0.25])
    gp, _ = O.gauss_volume(8)
    for q in range(8):
        _, sg, *_ = O.eval_shape_fem(8, gp[q], xl, np.zeros_like(xl))
        assert np.allclose(sg.T @ f, [1.5, -2.0, 0.25], atol=1e-12)


def test_pattern_closed_form_and_brute_force(oracle_mod):
    for dims in [(1, 1, 1), (3, 2, 4), (5, 4, 3)]:
        m = M.structured_hex(*dims)
        col, row = oracle_mod.sparsity(m.nnode, m.ien)
        n = [d + 1 for d in dims]
        assert row.size == np.prod([3 * k - 2 for k in n])
        # brute force set construction
        nb = [set() for _ in range(m.nnode)]
        for e in m.ien:
            for a in e:
                nb[a].update(e.tolist())
        for i in range(m.nnode):
            assert row[col[i]:col[i + 1]].tolist() == sorted(nb[i])


def test_spmv_oracle_equals_dense_product(oracle_mod):
    O = oracle_mod
    m = M.structured_hex(2, 2, 1)
    col, row = O.sparsity(m.nnode, m.ien)
    rng = np.random.default_rng(2)
    A = dict(k11=rng.normal(size=(row.size, 9)), g=rng.normal(size=(row.size, 3)),
             d1=rng.normal(size=(row.size, 3)), m=rng.normal(size=row.size))
    N = m.nnode
    D = np.zeros((4 * N, 4 * N))
    for i in range(N):
        for k in range(col[i], col[i + 1]):
            j = row[k]
            D[4 * i:4 * i + 3, 4 * j:4 * j + 3] = A["k11"][k].reshape(3, 3)
            D[4 * i:4 * i + 3, 4 * j + 3] = A["g"][k]
            D[4 * i + 3, 4 * j:4 * j + 3] = A["d1"][k]
            D[4 * i + 3, 4 * j + 3] = A["m"][k]
    xu = rng.normal(size=(N, 3))
    xp = rng.normal(size=N)
    yu, yp = O.spmv_up(col, row, A, xu, xp)
    y = D @ np.concatenate([xu, xp[:, None]], axis=1).ravel()
    assert np.allclose(np.concatenate([yu, yp[:, None]], axis=1).ravel(), y,
                       rtol=1e-13, atol=1e-13)


def test_gmres_oracle_against_direct_solve(oracle_mod):
    """Arnoldi with K >= n reproduces the direct solution of a small
    well-conditioned block system (block-Jacobi left preconditioning)."""
    O = oracle_mod
    m = M.structured_hex(1, 1, 2)
    col, row = O.sparsity(m.nnode, m.ien)
    N = m.nnode
    rng = np.random.default_rng(3)
    A = dict(k11=0.05 * rng.normal(size=(row.size, 9)), g=0.05 * rng.normal(size=(row.size, 3)),
             d1=0.05 * rng.normal(size=(row.size, 3)), m=0.05 * rng.normal(size=row.size))
    for i in range(N):
        k = col[i] + np.searchsorted(row[col[i]:col[i + 1]], i)
        A["k11"][k] += np.array([3, 0, 0, 0, 3, 0, 0, 0, 3.0])
        A["m"][k] += 3.0
    bu = rng.normal(size=(N, 3))
    bp = rng.normal(size=N)
    res = O.gmres_up(col, row, A, bu, bp, 1e-14, 4 * N, 0)
    yu, yp = O.spmv_up(col, row, A, res["solu"], res["solp"])
    r = np.concatenate([(yu - bu).ravel(), yp - bp])
    assert np.linalg.norm(r) < 1e-10 * np.linalg.norm(np.concatenate([bu.ravel(), bp]))


def test_genalpha_closed_form(oracle_mod):
    for rho in (0.0, 0.5, 1.0):
        ref = oracle_mod.genalpha(rho)
        g = C.genalpha(rho)
        assert np.array_equal(ref, [g["almi"], g["alfi"], g["gami"], g["beti"]])
    # rho_inf = 1: midpoint (almi = alfi = gami = 1/2)
    g = C.genalpha(1.0)
    assert g["almi"] == 0.5 and g["alfi"] == 0.5 and g["gami"] == 0.5


def test_channel_perturbation_is_divergence_free():
    x = np.random.default_rng(0).uniform(size=(50, 3)) * [4, 1, 1]
    h = 1e-6
    div = np.zeros(50)
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        up = C.div_free_perturbation(x + e, (4, 1, 1), seed=3, raw=True)
        um = C.div_free_perturbation(x - e, (4, 1, 1), seed=3, raw=True)
        div += (up[:, d] - um[:, d]) / (2 * h)
    u = C.div_free_perturbation(x, (4, 1, 1), seed=3, raw=True)
    assert np.abs(div).max() < 1e-6 * max(1.0, np.abs(u).max() * 20)


def test_problem_setups():
    p = C.cavity(4)
    assert p.mesh.nnode == 125 and p.mesh.nelem == 64
    top = np.abs(p.mesh.xg[:, 2] - 1.0) < 1e-12
    assert np.all(p.fields["ug"][top, 0] == 1.0)
    q = C.channel(8, 4, 4)
    inflow = np.abs(q.mesh.xg[:, 0]) < 1e-12
    assert np.all(q.ibc[inflow, :3] == 1)
    assert abs(q.fields["ug"][:, 0].max() - 1.0) < 0.05


def _oracle_lumped(O, mesh):
    """Row sums of the mass matrix from the oracle's shape functions and
    rule (eval_shape_fem, genGPandGW): sum_e sum_q N_a(q) DetJ(q) w_q."""
    nshl = mesh.ien.shape[1]
    gp, gw = O.gauss_volume(nshl)
    vol = np.zeros(mesh.nnode)
    for e in range(mesh.nelem):
        xl = mesh.xg[mesh.ien[e]]
        for q in range(len(gw)):
            shl, _, _, _, _, dj = O.eval_shape_fem(nshl, gp[q], xl, np.zeros_like(xl))
            vol[mesh.ien[e]] += shl * dj * gw[q]
    return vol


@pytest.mark.parametrize("split", ["hex", "tet", "prism"])
def test_lumped_volumes(oracle_mod, split):
    """The nodal volumes the heat-loss objective uses: on every element type
    they sum to the volume of the (jittered) box, and they equal the
    oracle's quadrature of the shape functions node by node."""
    m = M.structured_hex(4, 3, 3, hi=(1.0, 0.9, 1.1), jitter=0.08, seed=5)
    m = {"hex": m, "tet": M.hex_to_tets(m), "prism": M.hex_to_prisms(m)}[split]
    v = C.lumped_volumes(m)
    assert abs(v.sum() - 0.99) <= 1e-13
    w = _oracle_lumped(oracle_mod, m)
    assert np.abs(v - w).max() <= 1e-14 * np.abs(w).max()
    # a subset of the elements: the same quadrature over those alone
    sel = np.arange(m.nelem) % 3 == 0
    sub = M.Mesh(xg=m.xg, ien=m.ien[sel], elm_id=m.elm_id[sel])
    assert np.abs(C.lumped_volumes(m, sel) - _oracle_lumped(oracle_mod, sub)).max() <= \
        1e-14 * np.abs(w).max()
