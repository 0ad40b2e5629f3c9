"""Adjoint of the quenching temperature equation on the GPU (no reference
code).

* The per-node conductivity scale in the temperature assembly against the
  oracle (normwise 1e-12), discontinuity capturing on and off.
* LHSTem^T products and solves against the transposed sparse matrix.
* The element VJP of the residual (IntElmAss_Tem_Quenching_STAB as
  assembled) with respect to the alpha-stage T, rT, u and the conductivity
  scale by central differences (eps 1e-6, 16 seeded directions), for two
  materials: one with every property of order one, and the quench material
  (water and steel), where rho cp = 4.2e6 against a conductivity of 0.6 puts
  the conduction terms about eight orders below the residual whose rounding
  the differences see: there the tolerance (5e-5) sits at that rounding
  floor, ~1e-16 |lambda| |R| / eps over the derivative; the order-one case
  is held to 1e-6.  The LHS check below pins the T and rT derivatives to
  rounding.
* J_T,bc^T by central differences along Newton increments, its Dirichlet
  identity part, and the "LHS consistency gap" |LHSTem^T lam - J^T lam| /
  |J^T lam|: zero to rounding without discontinuity capturing (the LHS is
  then the exact Jacobian), nonzero with it.
* The adjoint solve J_T,bc^T lam = g (flexible GMRES preconditioned by
  LHSTem^T) to a 1e-8 true residual, FP64 and mixed.
* The VJP, J_T,bc^T and the adjoint solve again on linear tets and linear
  prisms (each hex of the tank split into 6 or 2), where both operators
  come from the dual-number kernel.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from oversetdevfort_amd import solver
    return solver


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# every property of order one (latent heat and saturation temperature
# included), so that no term of the temperature residual, and no
# temperature, is lost below the rounding of the others
UNIT_MATERIAL = dict(rhoa=0.5, rhow=1.0, cpa=0.8, cpw=1.0, kappaa=0.05, kappaw=0.2, rhos=2.0,
                     cps=0.5, kappas=1.0, lh=20.0, Ts=1.0, tref=1.0)


def _case(S, n=5, seed=1, tem_kdc=0.0, frac_ibc=0.15, kscale=True, unit=False, tet=False,
          prism=False):
    """A small quench tank with a perturbed state (temperatures on both sides
    of T_sat, VOF values in the interface range, random T Dirichlet nodes) and
    a per-node conductivity scale; ``unit``: the order-one material and a
    10 ms step; ``tet``, ``prism``: each hex split into 6 linear tets or 2
    linear prisms (the operators then come from the dual-number kernel)."""
    rng = np.random.default_rng(seed)
    prob = C.quench(n, seed=seed, jitter=0.05)
    if tet:
        prob.mesh = M.hex_to_tets(prob.mesh)
    if prism:
        prob.mesh = M.hex_to_prisms(prob.mesh)
    if unit:
        prob.material.update(UNIT_MATERIAL)
        prob.params["delt"] = 1e-2
    m, p = prob.mesh, prob.params
    N = m.nnode
    f = prob.fields
    f["ugold"] = 0.3 * rng.normal(size=(N, 3))
    f["ug"] = f["ugold"] + 0.05 * rng.normal(size=(N, 3))
    f["phigold"] = rng.uniform(0.0, 1.0, size=N)
    f["phig"] = f["phigold"] + 0.02 * rng.normal(size=N)
    Ts = prob.material["Ts"]                  # both sides of T_sat
    f["Tgold"] = Ts * rng.uniform(0.86, 1.15, size=N)
    f["Tg"] = f["Tgold"] + Ts * 2.5e-3 * rng.normal(size=N)
    f["rTgold"] = Ts * 0.05 * rng.normal(size=N)
    f["rTg"] = f["rTgold"] + Ts * 0.0125 * rng.normal(size=N)
    prob.ibc[:, 5] = (rng.uniform(size=N) < frac_ibc).astype(np.int32)
    mat = dict(prob.material, tem_kdc=tem_kdc)
    ctx = S.setup(prob)
    ctx.set_quench_params(mat)
    ks = 1.0 + 0.3 * rng.uniform(size=N) if kscale else None
    if ks is not None:
        ctx.set_field("KSCALE", ks)
    return ctx, prob, rng, mat, ks


def _rhs(ctx):
    ctx.assemble_tem(2)
    return ctx.get_tem_system()[1]


def _csr(O, prob, lhs):
    N = prob.mesh.nnode
    col, row = O.sparsity(N, prob.mesh.ien)
    return sp.csr_matrix((lhs, row, col), shape=(N, N))


@pytest.mark.parametrize("kdc", [0.0, 0.3])
def test_kscale_assembly_matches_oracle(S, oracle_mod, kdc):
    O = oracle_mod
    ctx, prob, rng, mat, ks = _case(S, seed=21, tem_kdc=kdc)
    m, p, f = prob.mesh, prob.params, prob.fields
    af, am = p["alfi"], p["almi"]
    st = dict(dg=f["dgold"] + af * (f["dg"] - f["dgold"]),
              ug=f["ugold"] + af * (f["ug"] - f["ugold"]),
              ugm=f["ugmold"] + af * (f["ugm"] - f["ugmold"]),
              phig=f["phigold"] + af * (f["phig"] - f["phigold"]),
              Tg=f["Tgold"] + af * (f["Tg"] - f["Tgold"]),
              rTg=f["rTgold"] + am * (f["rTg"] - f["rTgold"]))
    col, row = O.sparsity(m.nnode, m.ien)
    qv = [mat[k] for k in L.QPARAMS]
    R = O.assemble_tem(m, prob.ibc[:, 5], st, O.params_array(p), qv, col, row, 6, ks=ks)
    R1 = O.assemble_tem(m, prob.ibc[:, 5], st, O.params_array(p), qv, col, row, 6)
    ctx.assemble_tem(6)
    lhs, rhs = ctx.get_tem_system()
    ctx.set_field("KSCALE", np.ones(m.nnode))
    ctx.assemble_tem(6)
    lhs1, rhs1 = ctx.get_tem_system()
    ctx.close()
    assert _rel(lhs, R["lhs"]) <= 1e-12 and _rel(rhs, R["rhs"]) <= 1e-12
    # the conductivity part is ~1e-8 of the system (rho cp of water against
    # kappa): compare what the scale changes on its own
    dl, dr = R["lhs"] - R1["lhs"], R["rhs"] - R1["rhs"]
    print(f"tem_kdc {kdc}: the scale changes LHSTem by {_rel(R['lhs'], R1['lhs']):.1e}, "
          f"RHSGTem by {_rel(R['rhs'], R1['rhs']):.1e} (relative)", flush=True)
    assert _rel(lhs - lhs1, dl) <= 1e-6 and _rel(rhs - rhs1, dr) <= 1e-6


def test_tem_transposed_spmv_and_solve(S, oracle_mod):
    ctx, prob, rng, mat, ks = _case(S, n=6, seed=22)
    N = prob.mesh.nnode
    ctx.assemble_tem(6)
    lhs, _ = ctx.get_tem_system()
    T = _csr(oracle_mod, prob, lhs).T.tocsr()
    x = rng.normal(size=N)
    for prec, tol in ((L.OFD_FP64, 1e-13), (L.OFD_FP32, 1e-6)):
        assert _rel(ctx.SparseProd_LS_T(x, prec), T @ x) <= tol
    g = rng.normal(size=N) * (prob.ibc[:, 5] == 0)
    for prec in (L.OFD_FP64, L.OFD_FP32):
        res = ctx.solve_tem_t(g, rtol=1e-8, precision=prec)
        print(f"LHSTem^T solve precision {prec}: {res['cycles']} cycles, "
              f"{res['iters']} iterations", flush=True)
        assert np.linalg.norm(T @ res["lam"] - g) <= 1.01e-8 * np.linalg.norm(g)
    # the forward product is untouched by the transposed solve
    assert _rel(ctx.SparseProd_LS(x), T.T @ x) <= 1e-13
    ctx.close()


@pytest.mark.parametrize("kdc,unit,elem", [(0.0, True, "hex"), (0.3, True, "hex"),
                                           (0.0, False, "hex"), (0.3, False, "hex"),
                                           (0.3, True, "tet"), (0.0, False, "tet"),
                                           (0.0, True, "prism"), (0.3, False, "prism")])
def test_tem_vjp_matches_central_differences(S, kdc, unit, elem):
    ctx, prob, rng, mat, ks = _case(S, seed=23, tem_kdc=kdc, unit=unit, tet=elem == "tet",
                                    prism=elem == "prism")
    tol = 1e-6 if unit else 5e-5
    what = ("order-one material" if unit else "quench material") + f", {elem}"
    N = prob.mesh.nnode
    p, f = prob.params, prob.fields
    free = (prob.ibc[:, 5] == 0).astype(np.float64)
    lam = rng.normal(size=N) * free
    g = ctx.vjp_tem(lam)
    eps = 1e-6
    fields = (("TG", f["Tg"], p["alfi"] * g["T"], (N,)),
              ("RTG", f["rTg"], p["almi"] * g["rT"], (N,)),
              ("UG", f["ug"], p["alfi"] * g["u"].ravel(), (N, 3)),
              ("KSCALE", ks, g["ks"], (N,)))
    for name, base, grad, shape in fields:
        worst = 0.0
        for d in range(16):
            v = np.random.default_rng(800 + d).normal(size=shape)
            ctx.set_field(name, base + eps * v)
            fp = _rhs(ctx)
            ctx.set_field(name, base - eps * v)
            fm = _rhs(ctx)
            fd = lam @ (fp - fm) / (2 * eps)
            ad = grad @ v.ravel()
            worst = max(worst, abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
        ctx.set_field(name, base)
        print(f"{what}, tem_kdc {kdc}: VJP wrt {name}: worst relative FD error over 16 "
              f"directions {worst:.2e}", flush=True)
        assert worst <= tol, name
    ctx.close()


def _increment(ctx, f, v, s, p):
    """Newton increment s*v applied as the update does (rTg += d, Tg +=
    gami Delt d)."""
    ctx.set_field("RTG", f["rTg"] + s * v)
    ctx.set_field("TG", f["Tg"] + p["gami"] * p["delt"] * s * v)


@pytest.mark.parametrize("kdc,unit,elem", [(0.0, True, "hex"), (0.3, True, "hex"),
                                           (0.0, False, "hex"), (0.3, False, "hex"),
                                           (0.0, True, "tet"), (0.3, False, "tet"),
                                           (0.0, True, "prism")])
def test_tem_jt_apply_and_consistency_gap(S, kdc, unit, elem):
    ctx, prob, rng, mat, ks = _case(S, seed=24, tem_kdc=kdc, unit=unit, tet=elem == "tet",
                                    prism=elem == "prism")
    tol = 1e-6 if unit else 5e-5
    what = ("order-one material" if unit else "quench material") + f", {elem}"
    N = prob.mesh.nnode
    p, f = prob.params, prob.fields
    free = (prob.ibc[:, 5] == 0).astype(np.float64)
    lam = rng.normal(size=N) * free
    y = ctx.jt_apply_tem(lam)
    eps = 1e-6
    worst = 0.0
    for d in range(16):
        v = np.random.default_rng(900 + d).normal(size=N) * free
        _increment(ctx, f, v, eps, p)
        fp = _rhs(ctx)
        _increment(ctx, f, v, -eps, p)
        fm = _rhs(ctx)
        fd = lam @ (-(fp - fm) / (2 * eps))        # J = -dR/d(increment)
        ad = y @ v
        worst = max(worst, abs(fd - ad) / max(abs(fd), abs(ad)))
    _increment(ctx, f, np.zeros(N), 0.0, p)
    print(f"{what}, tem_kdc {kdc}: J_T^T worst relative FD error over 16 directions "
          f"{worst:.2e}", flush=True)
    assert worst <= tol
    # Dirichlet part: identity once per element sharing the node
    lam_d = rng.normal(size=N) * (1.0 - free)
    count = np.bincount(prob.mesh.ien.ravel(), minlength=N).astype(np.float64)
    assert np.allclose(ctx.jt_apply_tem(lam_d), count * lam_d, rtol=1e-15, atol=0)
    # LHS consistency gap
    ctx.assemble_tem(6)
    gap = _rel(ctx.SparseProd_LS_T(lam), y)
    print(f"{what}, tem_kdc {kdc}: LHS consistency gap |LHSTem^T lam - J_T^T lam| / "
          f"|J_T^T lam| = {gap:.3e}", flush=True)
    if kdc == 0.0:
        assert gap <= 1e-12
    else:
        assert gap > 1e-6
    ctx.close()


@pytest.mark.parametrize("elem", ["hex", "tet", "prism"])
def test_tem_adjoint_solve(S, elem):
    ctx, prob, rng, mat, ks = _case(S, n=6, seed=25, tem_kdc=0.3, tet=elem == "tet",
                                    prism=elem == "prism")
    N = prob.mesh.nnode
    ctx.assemble_tem(6)
    g = rng.normal(size=N) * (prob.ibc[:, 5] == 0)
    for prec in (L.OFD_FP64, L.OFD_FP32):
        res = ctx.solve_tem_adjoint(g, rtol=1e-8, precision=prec)
        print(f"temperature adjoint precision {prec}: {res['cycles']} restarts, "
              f"{res['iters']} inner iterations, relres {res['relres']:.2e}", flush=True)
        r = ctx.jt_apply_tem(res["lam"]) - g
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(g)
    ctx.close()
