"""The Fortran driver (ISO_C_BINDING -> libofd.so) against the Python
binding of the same C ABI: the drop-in SparseProdUP_3D / SparseGMRES_up
wrappers, two generalized-alpha steps of the cavity, and the adjoint
entries (flexible-GMRES adjoint solve, element VJP, transposed product) on
the final state must give the same numbers (both run the same
deterministic kernels in the same order)."""
import os
import subprocess

import numpy as np
import pytest

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(L.LIB_PATH), "fortran", "ofd_cavity_driver")


def _parse(out):
    rec = {"NEWTON": []}
    for line in out.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "NEWTON":
            rec["NEWTON"].append([float(v) for v in f[3:5]])
        elif f[0] in ("SPMVNORM", "GMRESNORM", "FIELDNORMS", "SPMVTNORM"):
            rec[f[0]] = [float(v) for v in f[1:]]
        elif f[0] == "ADJOINT":
            rec[f[0]] = [float(v) for v in f[1:6]]
    return rec


def _python_path(n, nsteps, nnewton):
    from oversetdevfort_amd import solver as S
    prob = C.cavity(n, jitter=0.0)
    ctx = S.setup(prob)
    rec = {"NEWTON": []}
    ctx.predictor()
    ctx.apply_dirichlet()
    ctx.assembleNavSto(6, 1)
    A = ctx.get_lhs()
    ru, rp = ctx.get_rhs()
    ctx.set_lhs(A)
    yu, yp = ctx.SparseProdUP_3D(ru, rp)
    rec["SPMVNORM"] = [np.linalg.norm(yu), np.linalg.norm(yp)]
    ctx.set_rhs(ru, rp)
    g = ctx.SparseGMRES_up(1e-6, 20, 2, L.OFD_FP64, L.OFD_ORTHO_MGS)
    rec["GMRESNORM"] = [np.linalg.norm(g["solu"]), np.linalg.norm(g["solp"])]
    ctx.set_field("UG", prob.fields["ug"])
    ctx.set_field("ACG", np.zeros((prob.mesh.nnode, 3)))
    ctx.set_field("PG", np.zeros(prob.mesh.nnode))
    for _ in range(nsteps):
        ctx.predictor()
        ctx.apply_dirichlet()
        for _ in range(nnewton):
            ctx.assembleNavSto(6, 1)
            ctx.solve(rtol=1e-10, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                      inner_tol=1e-4, kspace=30, max_cycles=100)
            ctx.newton_update()
            ctx.assembleNavSto(2, 1)
            rec["NEWTON"].append(list(ctx.rhs_norms()))
        ctx.advance()
    ug = ctx.get_field("UG")
    rec["FIELDNORMS"] = [np.linalg.norm(ug), np.linalg.norm(ctx.get_field("ACG")),
                         np.linalg.norm(ctx.get_field("PG"))]
    # the driver's adjoint section: velocity load on the free dofs, fine-scale
    # residual form, flexible GMRES, element VJP, transposed product
    ctx.set_residual_form(1)
    ctx.assembleNavSto(6, 1)
    su = ug * (prob.ibc[:, :3] == 0)
    sp = np.zeros(prob.mesh.nnode)
    res = ctx.solve_adjoint(su, sp, rtol=1e-10, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                            inner_tol=1e-3, kspace=30, max_cycles=20)
    g = ctx.vjp_ns(res["lamu"], res["lamp"])
    rec["ADJOINT"] = [np.linalg.norm(res["lamu"]), np.linalg.norm(res["lamp"]),
                      np.linalg.norm(g["u"]), np.linalg.norm(g["ac"]), np.linalg.norm(g["p"])]
    yu, yp = ctx.SparseProdUP_3D_T(su, res["lamp"])
    rec["SPMVTNORM"] = [np.linalg.norm(yu), np.linalg.norm(yp)]
    ctx.close()
    return rec


def test_fortran_driver_matches_python_binding():
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (make -C oversetdevfort_amd/fortran)")
    n, nsteps, nnewton = 6, 2, 2
    print(f"fortran driver cavity n={n} steps={nsteps} newton={nnewton}", flush=True)
    p = subprocess.run([EXE, str(n), str(nsteps), str(nnewton), "0"], capture_output=True,
                       text=True, timeout=100)
    assert p.returncode == 0, p.stdout + p.stderr
    f = _parse(p.stdout)
    py = _python_path(n, nsteps, nnewton)
    for key in ("SPMVNORM", "GMRESNORM", "FIELDNORMS", "ADJOINT", "SPMVTNORM"):
        np.testing.assert_allclose(f[key], py[key], rtol=1e-13, atol=0, err_msg=key)
    assert len(f["NEWTON"]) == nsteps * nnewton
    np.testing.assert_allclose(np.array(f["NEWTON"]), np.array(py["NEWTON"]), rtol=1e-12,
                               atol=1e-300)
