"""torch.autograd bindings of the hand-written adjoint kernels (the reference
has no adjoint).

Three functions sit behind ``torch.autograd.Function``:

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
* ``TemStep``: one time step of the quenching temperature equation with the
  flow held, (Tgold, rTgold, kscale, u) -> (Tg, rTg): predictor, then Newton
  iterations on IntElmAss_Tem_Quenching_STAB (solveQuenching.f90:579-882)
  with the assembled LHSTem (solveFlow.f90:139-140 updates) until the
  residual has dropped by ``rtol``.  Backward: one adjoint solve on the exact
  Jacobian J_T^T (flexible GMRES preconditioned by LHSTem^T) and one element
  VJP.

``run_steps`` / ``run_tem_steps`` chain the steps over time and, with
``checkpoint_every`` = k, keep only every k-th state for the reverse sweep:
each segment of k steps is re-solved during backward
(torch.utils.checkpoint).  The kernels are deterministic, so a re-solved
segment repeats the forward bit for bit.

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
           "run_steps", "TemStepper", "TemStep", "tem_step", "run_tem_steps",
           "solid_heat_loss"]


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


class TemStepper:
    """Solver settings of the quenching temperature equation on a
    ``solver.Context`` holding the mesh, pattern, parameters and quench
    material.  The flow is held: the VOF field ``phi`` (N,) is the same at the
    old and new level, the velocity is an input of each step.  ``ibc`` (N, 6+)
    Dirichlet flags, column 6 the temperature (the node keeps its old value).
    With TEM_KDC = 0 and C_DMDOT = 1 the assembled LHSTem is the exact
    Jacobian and the Newton iteration converges quadratically."""

    def __init__(self, ctx, ibc, params, phi, rtol=1e-12, max_newton=30, lin_rtol=1e-13,
                 adj_rtol=1e-12, precision=L.OFD_FP64, adj_precision=None, inner_tol=1e-3,
                 kspace=30, max_cycles=20):
        self.ctx = ctx
        self.free = (np.asarray(ibc)[:, 5] != 1).astype(np.float64)
        self.p = dict(params)
        self.phi = np.asarray(phi, np.float64)
        self.rtol, self.max_newton, self.lin_rtol = rtol, max_newton, lin_rtol
        self.adj_rtol, self.precision = adj_rtol, precision
        # inner solves of the adjoint: FP64 or mixed (FP32-stored LHSTem^T)
        self.adj_precision = precision if adj_precision is None else adj_precision
        self.inner_tol, self.kspace, self.max_cycles = inner_tol, kspace, max_cycles
        self.newton_history = []

    def _flow(self, u, ks):
        c = self.ctx
        c.set_field("UG", u)
        c.set_field("UGOLD", u)
        c.set_field("PHIG", self.phi)
        c.set_field("PHIGOLD", self.phi)
        c.set_field("KSCALE", ks)

    def solve(self, Told, rTold, ks, u):
        """One converged step from the old temperature; returns (Tg, rTg)."""
        c = self.ctx
        self._flow(u, ks)
        c.set_field("TGOLD", Told)
        c.set_field("RTGOLD", rTold)
        c.predictor()
        c.assemble_tem(2)
        r0 = r = c.tem_rhs_norm()
        hist = [r0]
        it = 0
        while r > self.rtol * r0:
            if it == self.max_newton:
                raise RuntimeError(f"temperature Newton did not converge: residual {hist}")
            c.assemble_tem(6)
            c.solve_tem(rtol=self.lin_rtol, precision=self.precision, kspace=self.kspace)
            c.tem_update()
            c.assemble_tem(2)
            r = c.tem_rhs_norm()
            hist.append(r)
            it += 1
        self.newton_history.append(hist)
        return c.get_field("TG"), c.get_field("RTG")

    def load(self, Told, rTold, ks, u, T, rT):
        c = self.ctx
        self._flow(u, ks)
        for name, v in (("TGOLD", Told), ("RTGOLD", rTold), ("TG", T), ("RTG", rT)):
            c.set_field(name, v)

    def step_vjp(self, gT, grT):
        """Reverse of ``solve`` on the loaded converged state: gradients with
        respect to (Tgold, rTgold, kscale, u) of <gT, Tg> + <grT, rTg>."""
        c, p = self.ctx, self.p
        ga, dt, af, am = p["gami"], p["delt"], p["alfi"], p["almi"]
        fr = self.free
        di = 1.0 - fr
        # seeds on the Newton increments (rTg += d, Tg += gami Delt d)
        s = (grT + ga * dt * gT) * fr
        if np.any(s):
            c.assemble_tem(6)                 # LHSTem preconditions J_T^T
            res = c.solve_tem_adjoint(s, rtol=self.adj_rtol, precision=self.adj_precision,
                                      inner_tol=self.inner_tol, kspace=self.kspace,
                                      max_cycles=self.max_cycles)
            lam = res["lam"] * fr
        else:
            lam = np.zeros_like(s)
        g = c.vjp_tem(lam)
        GT, GrT = g["T"], g["rT"]
        # alpha stage: free  T_a = Told + alfi Delt ((1-gami) rTold + gami rTg),
        #                    rT_a = (1-almi) rTold + almi rTg;
        #              Dirichlet  T = Told, rTg = (gami-1)/gami rTold
        d_Told = gT + GT
        d_rTold = (fr * ((1.0 - ga) * dt * gT + af * (1.0 - ga) * dt * GT + (1.0 - am) * GrT)
                   + di * ((ga - 1.0) / ga * grT + (1.0 - am / ga) * GrT))
        return d_Told, d_rTold, g["ks"], g["u"]


class TemStep(torch.autograd.Function):
    """(Tgold, rTgold, kscale, u) -> (Tg, rTg) of one converged temperature
    step with the flow held."""

    @staticmethod
    def forward(fc, st, Told, rTold, ks, u):
        T, rT = st.solve(_np(Told), _np(rTold), _np(ks), _np(u))
        T, rT = _t(T, Told), _t(rT, Told)
        fc.st = st
        fc.save_for_backward(Told, rTold, ks, u, T, rT)
        return T, rT

    @staticmethod
    def backward(fc, gT, grT):
        st = fc.st
        saved = fc.saved_tensors
        st.load(*(_np(v) for v in saved))
        z = np.zeros(st.free.shape)
        grads = st.step_vjp(_np(gT) if gT is not None else z,
                            _np(grT) if grT is not None else z)
        like = saved[0]
        return (None,) + tuple(_t(g, like) for g in grads)


def tem_step(st, Told, rTold, ks, u):
    return TemStep.apply(st, Told, rTold, ks, u)


def _tem_segment(st, n, T, rT, ks, u):
    for _ in range(n):
        T, rT = TemStep.apply(st, T, rT, ks, u)
    return T, rT


def run_tem_steps(st, Told, rTold, ks, u, nsteps, checkpoint_every=0):
    """``nsteps`` converged temperature steps, checkpointed as ``run_steps``."""
    state = (Told, rTold)
    done = 0
    k = int(checkpoint_every)
    while done < nsteps:
        n = min(k, nsteps - done) if k > 0 else nsteps - done
        if k > 0 and torch.is_grad_enabled():
            state = checkpoint(_tem_segment, st, n, *state, ks, u, use_reentrant=False)
        else:
            state = _tem_segment(st, n, *state, ks, u)
        done += n
    return state


def solid_heat_loss(st, Told, rTold, ks, u, rho_cp, solid_volume, nsteps, checkpoint_every=0):
    """Heat the solid gives off over ``nsteps`` temperature steps,
    sum_i V_i rho_s c_s (T0_i - T_i), V_i the solid's lumped nodal volumes:
    by the solid's energy balance, the time-integrated heat flux across its
    surface."""
    T, _ = run_tem_steps(st, Told, rTold, ks, u, nsteps, checkpoint_every)
    return torch.sum(rho_cp * solid_volume * (Told - T))
