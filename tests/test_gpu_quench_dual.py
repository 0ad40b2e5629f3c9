"""Exact derivatives of the quenching residuals by the dual-number kernels
(no reference counterpart), on the GPU, against central differences of the
forward assembly (the restatement of solveQuenching.f90:122-573 and
:579-882 that the oracle pins) and against each other:

* J_bc v from the stored exact Jacobian (ofd_assemble_nsvof_jac, product by
  ofd_spmv_nsvof) equals -dR/d(increment) v by central differences of the
  forward NS + VOF assembly, on the free rows, for increments zero at the
  constrained dofs; the constrained rows are v times the element count;
* the raw VJP (ofd_vjp_nsvof) against central differences of lambda . R
  along random directions of each alpha-stage field;
* the stored Jacobian's transpose against the raw VJP (two kernel modes);
* the dot-product test <lambda, A v> = <A^T lambda, v> of ofd_spmv_nsvof_t,
  and ofd_solve_nsvof_t's solution against the transposed product;
* the temperature VJP by dual numbers against the hand-written one
  (ofd_vjp_tem) and, for the VOF field, against central differences.

States: a 5^3 tank (trilinear hexes, and the same tank with each hex split
into 6 linear tets or 2 linear prisms) with perturbed fields, the VOF field strictly inside
(0, 1) (the clamps' kinks are away from the differences) and temperatures on
both sides of T_sat, random Dirichlet nodes.  Differences with steps of 1e-6
of each field's scale agree to about 1e-7; the tests hold 1e-5.
"""
import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C
from oversetdevfort_amd import mesh as M

pytestmark = pytest.mark.gpu

FD_TOL = 1e-5


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.fixture(scope="module", params=["hex", "tet", "prism"])
def case(request):
    from oversetdevfort_amd import solver as S
    rng = np.random.default_rng(7)
    prob = C.quench(5, seed=3, jitter=0.05)
    if request.param == "tet":
        prob.mesh = M.hex_to_tets(prob.mesh)
    elif request.param == "prism":
        prob.mesh = M.hex_to_prisms(prob.mesh)
    m = prob.mesh
    N = m.nnode
    f = prob.fields
    for k, s in (("ug", 0.3), ("ugm", 0.05), ("acg", 2.0), ("dg", 1e-3)):
        f[k + "old"] = s * rng.normal(size=(N, 3))
        f[k] = f[k + "old"] + 0.2 * s * rng.normal(size=(N, 3))
    f["pg"] = 100.0 * rng.normal(size=N)
    f["phigold"] = rng.uniform(0.1, 0.9, size=N)
    f["phig"] = np.clip(f["phigold"] + 0.02 * rng.normal(size=N), 0.05, 0.95)
    f["rphigold"] = 5.0 * rng.normal(size=N)
    f["rphig"] = f["rphigold"] + rng.normal(size=N)
    f["Tgold"] = rng.uniform(300.0, 600.0, size=N)
    f["Tg"] = f["Tgold"] + rng.normal(size=N)
    f["rTgold"] = 10.0 * rng.normal(size=N)
    f["rTg"] = f["rTgold"] + rng.normal(size=N)
    ibc = prob.ibc
    for c in (0, 1, 2, 6, 7):
        ibc[:, c] |= (rng.uniform(size=N) < 0.1).astype(np.int32)
    touched = np.zeros(N, bool)
    touched[m.ien[m.elm_id == 102].ravel()] = True
    ibc[np.ix_(~touched, [0, 1, 2, 6, 7])] = 1
    ctx = S.setup(prob)
    ctx.set_quench_params(prob.material)
    ks = np.exp(0.1 * rng.normal(size=N))
    ctx.set_field("KSCALE", ks)
    yield ctx, prob, rng
    ctx.close()


def _fields(ctx, f):
    for k, name in (("ug", "UG"), ("acg", "ACG"), ("pg", "PG"), ("phig", "PHIG"),
                    ("rphig", "RPHIG"), ("Tg", "TG"), ("rTg", "RTG")):
        ctx.set_field(name, f[k])


def _rhs(ctx, f):
    _fields(ctx, f)
    ctx.assemble_nsvof(2)
    _, r = ctx.get_nsvof_system()
    return r["rhsu"], r["rhsp"], r["rhsls"]


def _free(prob):
    ibc = prob.ibc
    return ((ibc[:, :3] != 1).astype(float), (ibc[:, 6] != 1).astype(float),
            (ibc[:, 7] != 1).astype(float))


def test_exact_jacobian_matches_differences(case):
    ctx, prob, rng = case
    p, f = prob.params, prob.fields
    N = prob.mesh.nnode
    fu, fp, fl = _free(prob)
    gdt, af = p["gami"] * p["delt"], p["alfi"]
    _fields(ctx, f)
    ctx.assemble_nsvof(2)
    ctx.assemble_nsvof_jac()
    for trial in range(3):
        vu = rng.normal(size=(N, 3)) * fu
        vp = 100.0 * rng.normal(size=N) * fp
        vl = rng.normal(size=N) * fl
        _fields(ctx, f)
        ju, jp, jl = ctx.SparseProd_NS_conv(vu, vp, vl)
        eps = 1e-6
        out = []
        for sgn in (1.0, -1.0):
            g = {k: v.copy() for k, v in f.items()}
            g["acg"] = f["acg"] + sgn * eps * vu
            g["ug"] = f["ug"] + sgn * eps * gdt * vu
            g["pg"] = f["pg"] + sgn * eps * af * gdt * vp
            g["rphig"] = f["rphig"] + sgn * eps * vl
            g["phig"] = f["phig"] + sgn * eps * gdt * vl
            out.append(_rhs(ctx, g))
        du = -(out[0][0] - out[1][0]) / (2 * eps)
        dp = -(out[0][1] - out[1][1]) / (2 * eps)
        dl = -(out[0][2] - out[1][2]) / (2 * eps)
        e = max(_rel(ju * fu, du), _rel(jp * fp, dp), _rel(jl * fl, dl))
        print(f"J v vs differences, trial {trial}: {e:.2e} "
              f"(|Jv| {np.linalg.norm(ju):.3e} {np.linalg.norm(jp):.3e} "
              f"{np.linalg.norm(jl):.3e})", flush=True)
        assert e <= FD_TOL
        # constrained rows: the element count times v (v is 0 there)
        assert not np.any(ju * (1 - fu)) and not np.any(jp * (1 - fp))
        assert not np.any(jl * (1 - fl))


def test_raw_vjp_matches_differences(case):
    ctx, prob, rng = case
    p, f = prob.params, prob.fields
    N = prob.mesh.nnode
    fu, fp, fl = _free(prob)
    lu = rng.normal(size=(N, 3)) * fu
    lp = rng.normal(size=N) * fp
    ll = rng.normal(size=N) * fl
    _fields(ctx, f)
    g = ctx.vjp_nsvof(lu, lp, ll)
    # field -> (gradient key, alpha-stage factor, scale of the field)
    specs = {"ug": ("u", p["alfi"], 0.3), "acg": ("ac", p["almi"], 2.0),
             "pg": ("p", 1.0, 100.0), "phig": ("phi", p["alfi"], 0.1),
             "rphig": ("rphi", p["almi"], 5.0), "Tg": ("T", p["alfi"], 100.0)}
    for fld, (key, fac, scale) in specs.items():
        w = rng.normal(size=f[fld].shape)
        eps = 1e-6 * scale
        vals = []
        for sgn in (1.0, -1.0):
            h = {k: v.copy() for k, v in f.items()}
            h[fld] = f[fld] + sgn * eps * w
            ru, rp, rl = _rhs(ctx, h)
            vals.append(np.sum(lu * ru) + np.sum(lp * rp) + np.sum(ll * rl))
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = fac * np.sum(g[key] * w)
        e = abs(fd - ad) / max(abs(fd), 1e-300)
        print(f"VJP {key}: adjoint {ad:.10e} differences {fd:.10e} rel {e:.2e}", flush=True)
        assert e <= FD_TOL


def test_jacobian_transpose_matches_vjp(case):
    ctx, prob, rng = case
    p, f = prob.params, prob.fields
    N = prob.mesh.nnode
    fu, fp, fl = _free(prob)
    gdt, af, am = p["gami"] * p["delt"], p["alfi"], p["almi"]
    _fields(ctx, f)
    ctx.assemble_nsvof(2)
    ctx.assemble_nsvof_jac()
    lu = rng.normal(size=(N, 3)) * fu
    lp = rng.normal(size=N) * fp
    ll = rng.normal(size=N) * fl
    tu, tp, tl = ctx.SparseProd_NS_conv_T(lu, lp, ll)
    g = ctx.vjp_nsvof(lu, lp, ll)
    # J = -dR/d(increment): d moves ac by almi, u by alfi gami Delt; dp moves
    # p by alfi gami Delt; dphi moves rphi by almi, phi by alfi gami Delt
    ru = -(am * g["ac"] + af * gdt * g["u"])
    rp = -(af * gdt * g["p"])
    rl = -(am * g["rphi"] + af * gdt * g["phi"])
    e = max(_rel(tu * fu, ru * fu), _rel(tp * fp, rp * fp), _rel(tl * fl, rl * fl))
    print(f"stored J^T lambda vs raw VJP: {e:.2e}", flush=True)
    assert e <= 1e-12
    # dot-product test of the transposed product
    vu = rng.normal(size=(N, 3))
    vp = rng.normal(size=N)
    vl = rng.normal(size=N)
    au, ap, al = ctx.SparseProd_NS_conv(vu, vp, vl)
    lhs = np.sum(lu * au) + np.sum(lp * ap) + np.sum(ll * al)
    rhs = np.sum(tu * vu) + np.sum(tp * vp) + np.sum(tl * vl)
    e = abs(lhs - rhs) / abs(lhs)
    print(f"<lambda, A v> {lhs:.15e} <A^T lambda, v> {rhs:.15e} rel {e:.1e}", flush=True)
    assert e <= 1e-12


def test_transposed_solve(case):
    ctx, prob, rng = case
    f = prob.fields
    N = prob.mesh.nnode
    fu, fp, fl = _free(prob)
    _fields(ctx, f)
    ctx.assemble_nsvof(2)
    ctx.assemble_nsvof_jac()
    gu = rng.normal(size=(N, 3)) * fu
    gp = rng.normal(size=N) * fp
    gl = rng.normal(size=N) * fl
    for prec in (L.OFD_FP64, L.OFD_FP32):
        res = ctx.solve_nsvof_t(gu, gp, gl, rtol=1e-10, precision=prec)
        tu, tp, tl = ctx.SparseProd_NS_conv_T(res["lamu"], res["lamp"], res["lamphi"])
        e = np.sqrt(np.sum((tu - gu) ** 2) + np.sum((tp - gp) ** 2) + np.sum((tl - gl) ** 2))
        e /= np.sqrt(np.sum(gu ** 2) + np.sum(gp ** 2) + np.sum(gl ** 2))
        print(f"A^T lambda = g ({'fp64' if prec == L.OFD_FP64 else 'mixed'} inner): "
              f"{res['iters']} inner iterations, {res['cycles']} cycles, true residual "
              f"{e:.2e}", flush=True)
        assert res["converged"] and e <= 1e-9


def test_temperature_dual_vjp(case):
    ctx, prob, rng = case
    p, f = prob.params, prob.fields
    N = prob.mesh.nnode
    free = (prob.ibc[:, 5] != 1).astype(float)
    lam = rng.normal(size=N) * free
    _fields(ctx, f)
    gd = ctx.vjp_tem_dual(lam)
    gh = ctx.vjp_tem(lam)
    for k in ("T", "rT", "u", "ks"):
        e = _rel(gd[k], gh[k])
        print(f"T VJP {k}: dual vs hand-written {e:.2e}", flush=True)
        assert e <= 1e-12
    # the VOF field, which the hand-written VJP holds
    w = rng.normal(size=N)
    eps = 1e-7
    vals = []
    for sgn in (1.0, -1.0):
        h = {k: v.copy() for k, v in f.items()}
        h["phig"] = f["phig"] + sgn * eps * w
        _fields(ctx, h)
        ctx.assemble_tem(2)
        _, rT = ctx.get_tem_system()
        vals.append(np.sum(lam * rT))
    fd = (vals[0] - vals[1]) / (2 * eps)
    ad = p["alfi"] * np.sum(gd["phi"] * w)
    e = abs(fd - ad) / abs(fd)
    print(f"T VJP phi: adjoint {ad:.10e} differences {fd:.10e} rel {e:.2e}", flush=True)
    assert e <= FD_TOL
