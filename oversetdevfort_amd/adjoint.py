"""torch.autograd bindings of the hand-written adjoint kernels (the reference
has no adjoint).

Two functions sit behind ``torch.autograd.Function``:

* ``NSResidual``: the Navier-Stokes residual R(ug, acg, pg, muscale) as
  assembled by IntElmAss_NS_overset (assembleNavSto.f90:76-336) at the
  generalized-alpha stage of (ug, acg), with the Dirichlet rows of BCLhs_3D
  (BCLhs_3D.f90:4-122) zero.  Backward: the element VJP kernel.
* ``NavStoStep``: one time step of the NS block, (ugold, acgold, pgold,
  muscale, bcval) -> (ug, acg, pg): predictor (driver.f90:83-85), Dirichlet
  values with the consistent acceleration (setBCs.f90:84-91), then exact
  Newton iterations (flexible GMRES on the element Jacobian, preconditioned by
  the assembled LHS) until the residual has dropped by ``rtol``.  Backward:
  one adjoint solve J^T lam = s on the converged state and one element VJP,
  by the implicit function theorem; the step's explicit dependence on the old
  state follows the predictor/update formulas (Newmark in the increments).

``run_steps`` chains ``NavStoStep`` over time and, with ``checkpoint_every``
= k, keeps only every k-th state for the reverse sweep: each segment of k
steps is re-solved during backward (torch.utils.checkpoint).  The kernels
are deterministic, so a re-solved segment repeats the forward bit for bit.

Every array crosses the C ABI through host memory; tensors may live on any
device and gradients come back on the device of the input.  The solver
context holds one state at a time: forward and backward load what they need
into it, so several functions may share one context.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from . import _lib as L

__all__ = ["NSStepper", "NSResidual", "NavStoStep", "ns_residual", "nav_sto_step",
           "run_steps"]


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


def _t(a, like):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=like.device)


class NSStepper:
    """Solver settings of one NS block on a ``solver.Context``.

    ``ibc`` (NNODE, 4+) Dirichlet flags (1 = fixed) of u, v, w, p.  The
    context must hold the mesh, pattern and parameters; its residual form is
    set here (1: fine-scale velocity uprime = -tauM rLi, whose Newton
    iteration converges to machine precision; 0: the reference's uprime = 0,
    which has no exact root for enclosed flows)."""

    def __init__(self, ctx, ibc, params, residual_form=1, subdomain_flag=1, rtol=1e-12,
                 max_newton=12, lin_rtol=1e-12, adj_rtol=1e-12, precision=L.OFD_FP64,
                 inner_tol=1e-3, kspace=30, max_cycles=20):
        self.ctx = ctx
        b = np.asarray(ibc)
        self.free_u = (b[:, :3] != 1).astype(np.float64)
        self.free_p = (b[:, 3] != 1).astype(np.float64)
        self.p = dict(params)
        self.sub = int(subdomain_flag)
        self.rtol, self.max_newton, self.lin_rtol = rtol, max_newton, lin_rtol
        self.adj_rtol, self.precision = adj_rtol, precision
        self.inner_tol, self.kspace, self.max_cycles = inner_tol, kspace, max_cycles
        ctx.set_residual_form(residual_form)
        self.newton_history = []

    def _norm(self):
        ru, rp = self.ctx.rhs_norms()
        return float(np.hypot(ru, rp))

    # ---- forward ----
    def solve(self, ugold, acgold, pgold, mus, bcval):
        """One converged time step from the old state; returns (ug, acg, pg)."""
        c = self.ctx
        c.set_field("UGOLD", ugold)
        c.set_field("ACGOLD", acgold)
        c.set_field("PGOLD", pgold)
        c.set_field("MUSCALE", mus)
        c.set_bcvals(bcval)
        c.predictor()
        c.apply_dirichlet()
        c.assembleNavSto(2, self.sub)
        r0 = r = self._norm()
        hist = [r0]
        it = 0
        while r > self.rtol * r0:
            if it == self.max_newton:
                raise RuntimeError(f"Newton did not converge: residual {hist}")
            c.assembleNavSto(6, self.sub)
            c.solve_newton(rtol=self.lin_rtol, subdomain_flag=self.sub,
                           precision=self.precision, inner_tol=self.inner_tol,
                           kspace=self.kspace, max_cycles=self.max_cycles)
            c.newton_update()
            c.assembleNavSto(2, self.sub)
            r = self._norm()
            hist.append(r)
            it += 1
        self.newton_history.append(hist)
        return c.get_field("UG"), c.get_field("ACG"), c.get_field("PG")

    # ---- reverse ----
    def load(self, ugold, acgold, pgold, mus, bcval, ug, acg, pg):
        c = self.ctx
        for name, v in (("UGOLD", ugold), ("ACGOLD", acgold), ("PGOLD", pgold),
                        ("MUSCALE", mus), ("UG", ug), ("ACG", acg), ("PG", pg)):
            c.set_field(name, v)
        c.set_bcvals(bcval)

    def step_vjp(self, gug, gacg, gpg):
        """Reverse of ``solve`` on the loaded converged state: gradients with
        respect to (ugold, acgold, pgold, mus, bcval) of
        <gug, ug> + <gacg, acg> + <gpg, pg>."""
        c, p = self.ctx, self.p
        ga, dt, af, am = p["gami"], p["delt"], p["alfi"], p["almi"]
        fu, fp = self.free_u, self.free_p
        du, dp = 1.0 - fu, 1.0 - fp
        gdt = ga * dt
        # seeds on the Newton increments d (acg += d, ug += gami Delt d,
        # pg += alfi gami Delt dp) of the free dofs
        su = (gacg + gdt * gug) * fu
        sp = af * gdt * gpg * fp
        if np.any(su) or np.any(sp):
            c.assembleNavSto(6, self.sub)      # the LHS that preconditions J^T
            res = c.solve_adjoint(su, sp, rtol=self.adj_rtol, subdomain_flag=self.sub,
                                  precision=self.precision, inner_tol=self.inner_tol,
                                  kspace=self.kspace, max_cycles=self.max_cycles)
            lu, lp = res["lamu"] * fu, res["lamp"] * fp
        else:
            lu, lp = np.zeros_like(su), np.zeros_like(sp)
        g = c.vjp_ns(lu, lp, self.sub)        # d(lam . R) / d(alpha-stage u, ac, p, mus)
        Gu, Gac, Gp = g["u"], g["ac"], g["p"]
        # alpha stage: free  u_a = uold + alfi (1-gami) Delt acold + alfi gami Delt acg,
        #                    ac_a = (1-almi) acold + almi acg;
        #              Dirichlet  u_a = (1-alfi) uold + alfi g,
        #                    acg = (g - uold) / (gami Delt) + (gami-1)/gami acold
        d_ugold = fu * (gug + Gu) + du * ((1.0 - af) * Gu - am / gdt * Gac - gacg / gdt)
        d_acgold = (fu * ((1.0 - ga) * dt * gug + af * (1.0 - ga) * dt * Gu + (1.0 - am) * Gac)
                    + du * (((1.0 - am) - am * (1.0 - ga) / ga) * Gac - (1.0 - ga) / ga * gacg))
        d_pgold = dp * (gpg + Gp)
        d_bcval = du * (gug + gacg / gdt + af * Gu + am / gdt * Gac)
        return d_ugold, d_acgold, d_pgold, g["mus"], d_bcval


class NSResidual(torch.autograd.Function):
    """(ug, acg, pg, muscale) -> (Ru (N,3), Rp (N,)), the old state taken
    from the context at call time."""

    @staticmethod
    def forward(fc, st, ug, acg, pg, mus):
        c = st.ctx
        old = [torch.as_tensor(c.get_field(n)) for n in ("UGOLD", "ACGOLD", "PGOLD")]
        for name, v in (("UG", ug), ("ACG", acg), ("PG", pg), ("MUSCALE", mus)):
            c.set_field(name, _np(v))
        c.assembleNavSto(2, st.sub)
        ru, rp = c.get_rhs()
        fc.st = st
        fc.save_for_backward(ug, acg, pg, mus, *old)
        return _t(ru, ug), _t(rp, ug)

    @staticmethod
    def backward(fc, gru, grp):
        st = fc.st
        ug, acg, pg, mus, ugold, acgold, pgold = fc.saved_tensors
        c, p = st.ctx, st.p
        for name, v in (("UGOLD", ugold), ("ACGOLD", acgold), ("PGOLD", pgold), ("UG", ug),
                        ("ACG", acg), ("PG", pg), ("MUSCALE", mus)):
            c.set_field(name, _np(v))
        lu = (_np(gru) if gru is not None else np.zeros(st.free_u.shape)) * st.free_u
        lp = (_np(grp) if grp is not None else np.zeros(st.free_p.shape)) * st.free_p
        g = c.vjp_ns(lu, lp, st.sub)
        return (None, _t(p["alfi"] * g["u"], ug), _t(p["almi"] * g["ac"], ug),
                _t(g["p"], ug), _t(g["mus"], ug))


class NavStoStep(torch.autograd.Function):
    """(ugold, acgold, pgold, muscale, bcval) -> (ug, acg, pg) of one
    converged time step."""

    @staticmethod
    def forward(fc, st, ugold, acgold, pgold, mus, bcval):
        ug, acg, pg = st.solve(_np(ugold), _np(acgold), _np(pgold), _np(mus), _np(bcval))
        ug, acg, pg = _t(ug, ugold), _t(acg, ugold), _t(pg, ugold)
        fc.st = st
        fc.save_for_backward(ugold, acgold, pgold, mus, bcval, ug, acg, pg)
        return ug, acg, pg

    @staticmethod
    def backward(fc, gug, gacg, gpg):
        st = fc.st
        saved = fc.saved_tensors
        st.load(*(_np(v) for v in saved))
        z3 = np.zeros(st.free_u.shape)
        z1 = np.zeros(st.free_p.shape)
        grads = st.step_vjp(_np(gug) if gug is not None else z3,
                            _np(gacg) if gacg is not None else z3,
                            _np(gpg) if gpg is not None else z1)
        like = saved[0]
        return (None,) + tuple(_t(g, like) for g in grads)


def ns_residual(st, ug, acg, pg, mus):
    return NSResidual.apply(st, ug, acg, pg, mus)


def nav_sto_step(st, ugold, acgold, pgold, mus, bcval):
    return NavStoStep.apply(st, ugold, acgold, pgold, mus, bcval)


def _segment(st, n, ug, acg, pg, mus, bcval):
    for _ in range(n):
        ug, acg, pg = NavStoStep.apply(st, ug, acg, pg, mus, bcval)
    return ug, acg, pg


def run_steps(st, ugold, acgold, pgold, mus, bcval, nsteps, checkpoint_every=0):
    """``nsteps`` converged time steps; with ``checkpoint_every`` = k > 0 and
    autograd recording, only the state entering each k-step segment is kept
    and the segment is re-solved in the reverse sweep."""
    state = (ugold, acgold, pgold)
    done = 0
    k = int(checkpoint_every)
    while done < nsteps:
        n = min(k, nsteps - done) if k > 0 else nsteps - done
        if k > 0 and torch.is_grad_enabled():
            state = checkpoint(_segment, st, n, *state, mus, bcval, use_reentrant=False)
        else:
            state = _segment(st, n, *state, mus, bcval)
        done += n
    return state
