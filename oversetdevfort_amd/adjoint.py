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

* ``QuenchStep``: one coupled step of the quenching solver, (old u, ac, p,
  phi, rphi, T, rT, kscale, inflow amplitude) -> the new fields: the
  staggered iteration of solveFlow.f90 with exact Newton steps on both
  systems (the NS + VOF Jacobian by dual numbers).  Backward: the coupled
  transposed system by block Gauss-Seidel (one temperature adjoint solve
  and one NS + VOF transposed solve per sweep), then the residuals' VJPs.

``run_steps`` / ``run_tem_steps`` / ``run_quench_steps`` chain the steps over time and, with
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
           "solid_heat_loss", "QuenchStepper", "QuenchStep", "quench_step",
           "run_quench_steps", "quench_heat_loss"]


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


class QuenchStepper:
    """One time step of the quenching solver with the flow, the VOF field and
    the temperature coupled, differentiable end to end.

    Forward: predictor (driver.f90:83-85) and the Dirichlet values of the
    velocity, amp x ``profile`` on the inlet and 0 on the walls (setBCs_CFD,
    setBCs.f90:84-91), then the reference's staggered iteration
    (solmultiphasethermofluid_stag, solveFlow.f90:57-161) with exact Newton
    steps: per iteration the NS + VOF system with its exact Jacobian
    (ofd_assemble_nsvof_jac) and the temperature system with LHSTem (exact
    for TEM_KDC = 0 and C_DMDOT = 1), until every residual (u, p, phi, T) has
    dropped by ``rtol`` from the predictor's or sits at its rounding floor.

    Backward, at the converged state, by the implicit function theorem: the
    seeds on the new fields become seeds on the Newton increments, and the
    transposed coupled system
        [J_FF^T  J_TF^T] [lam_F]   [s_F]
        [J_FT^T  J_TT^T] [lam_T] = [s_T]
    is solved by block Gauss-Seidel, each sweep one temperature adjoint
    solve (ofd_solve_tem_adjoint) and one NS + VOF transposed solve
    (ofd_solve_nsvof_t on the exact Jacobian), the couplings
    J_FT^T lam_F = -alfi gami Delt dR_F/dT . lam_F and
    J_TF^T lam_T = -alfi gami Delt dR_T/d(u, phi) . lam_T from the dual-number
    VJPs.  Then one VJP of each residual gives the gradients with respect to
    the old state, the conductivity scale and the inflow amplitude.

    ``ibc`` (N, 8): columns 1-3 u, 6 T, 7 p, 8 phi (1 = fixed)."""

    NEW = ("UG", "ACG", "PG", "PHIG", "RPHIG", "TG", "RTG")
    OLD = ("UGOLD", "ACGOLD", "PGOLD", "PHIGOLD", "RPHIGOLD", "TGOLD", "RTGOLD")

    def __init__(self, ctx, ibc, params, profile, rtol=1e-12, max_newton=12, lin_rtol=1e-12,
                 adj_rtol=1e-11, precision=L.OFD_FP64, adj_precision=None, inner_tol=1e-3,
                 kspace=30, max_cycles=200, gs_rtol=1e-11, max_gs=40):
        b = np.asarray(ibc)
        self.ctx = ctx
        self.fu = (b[:, :3] != 1).astype(np.float64)
        self.fp = (b[:, 6] != 1).astype(np.float64)
        self.fphi = (b[:, 7] != 1).astype(np.float64)
        self.fT = (b[:, 5] != 1).astype(np.float64)
        self.p = dict(params)
        self.profile = np.asarray(profile, np.float64)
        self.rtol, self.max_newton, self.lin_rtol = rtol, max_newton, lin_rtol
        self.adj_rtol, self.precision = adj_rtol, precision
        self.adj_precision = precision if adj_precision is None else adj_precision
        self.inner_tol, self.kspace, self.max_cycles = inner_tol, kspace, max_cycles
        self.gs_rtol, self.max_gs = gs_rtol, max_gs
        self.newton_history = []
        self.gs_history = []

    def _load_old(self, old, ks, amp):
        c = self.ctx
        for name, v in zip(self.OLD, old):
            c.set_field(name, v)
        c.set_field("KSCALE", ks)
        c.set_bcvals(float(amp) * self.profile)

    def _norms(self):
        c = self.ctx
        c.assemble_nsvof(2)
        r = c.nsvof_rhs_norms()
        c.assemble_tem(2)
        return (*r, c.tem_rhs_norm())

    def _converged(self, hist):
        r0, prev, r = hist[0], hist[-2], hist[-1]
        # a component is done once it has dropped by rtol, or once it has
        # dropped by 1e-9 and stopped falling (its rounding floor)
        return all(ri <= self.rtol * r0i or (ri <= 1e-9 * r0i and ri > 0.2 * pi) or ri == 0.0
                   for ri, r0i, pi in zip(r, r0, prev))

    # ---- forward ----
    def solve(self, old, ks, amp):
        """One converged time step from the old state; returns the new
        (ug, acg, pg, phig, rphig, Tg, rTg)."""
        c = self.ctx
        self._load_old(old, ks, amp)
        c.predictor()
        c.apply_dirichlet()
        hist = [self._norms()]
        it = 0
        while it == 0 or not self._converged(hist):
            if it == self.max_newton:
                raise RuntimeError(f"coupled Newton did not converge: residuals {hist}")
            # NS + VOF: the residual assembled by _norms, the exact Jacobian
            # (a solve that stops at its rounding floor short of lin_rtol
            # still gives a Newton step; the residual test decides)
            c.assemble_nsvof_jac()
            c.solve_nsvof(rtol=self.lin_rtol, precision=self.precision,
                          ortho=L.OFD_ORTHO_CGS1 | L.OFD_PREC_RIGHT, kspace=self.kspace,
                          max_cycles=self.max_cycles, raise_noconv=False)
            c.nsvof_update()
            # temperature with the new flow
            c.assemble_tem(6)
            c.solve_tem(rtol=self.lin_rtol, precision=self.precision, kspace=self.kspace,
                        max_cycles=self.max_cycles, raise_noconv=False)
            c.tem_update()
            hist.append(self._norms())
            it += 1
        self.newton_history.append(hist)
        return tuple(c.get_field(n) for n in self.NEW)

    # ---- reverse ----
    @staticmethod
    def _checked(res, floor=1e-8):
        # an adjoint solve may stop at its rounding floor short of adj_rtol;
        # one that stays above `floor` is an error
        if not res["converged"] and not res["relres"] <= floor:
            raise RuntimeError(f"adjoint solve stopped at relres {res['relres']:.3e}")
        return res

    def _solve_T(self, rhs, lam0):
        """J_TT^T lam = rhs (free T dofs) from lam0: the correction for the
        residual rhs - J_TT^T lam0 (matrix-free), solved so that lam's
        residual is adj_rtol |rhs|.  Returns lam and its relative residual."""
        c = self.ctx
        nb = float(np.linalg.norm(rhs))
        if nb == 0.0:
            return np.zeros_like(rhs), 0.0
        r = rhs - c.jt_apply_tem(lam0) * self.fT if np.any(lam0) else rhs
        nr = float(np.linalg.norm(r))
        if nr <= self.adj_rtol * nb:
            return lam0, nr / nb
        res = self._checked(c.solve_tem_adjoint(
            r, rtol=min(0.1, self.adj_rtol * nb / nr), precision=self.adj_precision,
            inner_tol=self.inner_tol, kspace=self.kspace, max_cycles=self.max_cycles,
            raise_noconv=False), floor=1e-8 * nb / nr)
        return lam0 + res["lam"] * self.fT, res["relres"] * nr / nb

    def _solve_F(self, ru, rp, rl, lam0):
        """J_FF^T lam = (ru, rp, rl) on the stored exact Jacobian from lam0, as
        ``_solve_T``."""
        c = self.ctx
        fu, fp, fl = self.fu, self.fp, self.fphi
        nb = float(np.sqrt(np.sum(ru ** 2) + np.sum(rp ** 2) + np.sum(rl ** 2)))
        if nb == 0.0:
            return (np.zeros_like(ru), np.zeros_like(rp), np.zeros_like(rl)), 0.0
        if any(np.any(v) for v in lam0):
            tu, tp, tl = c.SparseProd_NS_conv_T(*lam0)
            ru, rp, rl = (ru - tu) * fu, (rp - tp) * fp, (rl - tl) * fl
        nr = float(np.sqrt(np.sum(ru ** 2) + np.sum(rp ** 2) + np.sum(rl ** 2)))
        if nr <= self.adj_rtol * nb:
            return lam0, nr / nb
        r = self._checked(c.solve_nsvof_t(
            ru, rp, rl, rtol=min(0.1, self.adj_rtol * nb / nr), precision=self.adj_precision,
            inner_tol=self.inner_tol, kspace=self.kspace, max_cycles=self.max_cycles,
            raise_noconv=False), floor=1e-8 * nb / nr)
        return ((lam0[0] + r["lamu"] * fu, lam0[1] + r["lamp"] * fp,
                 lam0[2] + r["lamphi"] * fl), r["relres"] * nr / nb)

    def load(self, old, ks, amp, new):
        self._load_old(old, ks, amp)
        for name, v in zip(self.NEW, new):
            self.ctx.set_field(name, v)

    def step_vjp(self, g):
        """Reverse of ``solve`` on the loaded converged state.  ``g``: seeds
        on the new fields (ug (N,3), acg (N,3), pg, phig, rphig, Tg, rTg).
        Returns the gradients with respect to the old fields (same order),
        the conductivity scale and the inflow amplitude."""
        c, p = self.ctx, self.p
        ga, dt, af, am = p["gami"], p["delt"], p["alfi"], p["almi"]
        gdt = ga * dt
        fu, fp, fl, fT = self.fu, self.fp, self.fphi, self.fT
        gug, gacg, gpg, gphig, grphig, gTg, grTg = g
        # seeds on the Newton increments of the free dofs (acg += d, ug +=
        # gami Delt d; pg += alfi gami Delt dp; rphig += dphi, phig += gami
        # Delt dphi; rTg += dT, Tg += gami Delt dT)
        su = (gacg + gdt * gug) * fu
        sp = af * gdt * gpg * fp
        sl = (grphig + gdt * gphig) * fl
        sT = (grTg + gdt * gTg) * fT
        N = len(sp)
        lu, lp, ll = np.zeros((N, 3)), np.zeros(N), np.zeros(N)
        lT = np.zeros(N)
        sweeps = []
        if np.any(su) or np.any(sp) or np.any(sl) or np.any(sT):
            c.assemble_tem(6)                  # LHSTem preconditions J_TT^T
            c.assemble_nsvof(2)
            c.assemble_nsvof_jac()             # J_FF, transposed in the solve
            for k in range(self.max_gs):
                gF = c.vjp_nsvof(lu, lp, ll)
                rT = (sT + af * gdt * gF["T"]) * fT
                lT, eT = self._solve_T(rT, lT)
                gT = c.vjp_tem_dual(lT)
                ru = (su + af * gdt * gT["u"]) * fu
                rl = (sl + af * gdt * gT["phi"]) * fl
                (nu, npp, nl), eF = self._solve_F(ru, sp, rl, (lu, lp, ll))
                dn = np.sqrt(np.sum((nu - lu) ** 2) + np.sum((npp - lp) ** 2) +
                             np.sum((nl - ll) ** 2))
                nn = np.sqrt(np.sum(nu ** 2) + np.sum(npp ** 2) + np.sum(nl ** 2))
                lu, lp, ll = nu, npp, nl
                sweeps.append(dn / max(nn, 1e-300))
                # done once a sweep changes lambda_F by less than gs_rtol, by
                # less than 100 times what the sweep's two solves left, or by
                # less than 1e-6 without falling further (the accuracy the
                # solves give lambda: their residual times the condition
                # number)
                floor = 100.0 * max(eT, eF)
                stalled = (len(sweeps) > 2 and sweeps[-1] <= 1e-6 and
                           sweeps[-1] >= 0.3 * sweeps[-2])
                if sweeps[-1] <= max(self.gs_rtol, floor) or stalled:
                    break
            else:
                raise RuntimeError(f"coupled adjoint sweeps did not converge: {sweeps}")
            # the last lam_T belongs to the lam_F before the last sweep: redo it
            gF = c.vjp_nsvof(lu, lp, ll)
            rT = (sT + af * gdt * gF["T"]) * fT
            lT, _ = self._solve_T(rT, lT)
        self.gs_history.append(sweeps)
        gF = c.vjp_nsvof(lu, lp, ll)
        gT = c.vjp_tem_dual(lT)
        Gu = gF["u"] + gT["u"]
        Gac, Gp, Grphi = gF["ac"], gF["p"], gF["rphi"]
        Gphi = gF["phi"] + gT["phi"]
        GT = gF["T"] + gT["T"]
        GrT, Gks = gT["rT"], gT["ks"]
        # explicit dependence at fixed increments (from the predictor):
        # free u: ug = ugold + gami Delt d, acg = (gami-1)/gami acgold + d;
        # Dirichlet u: ug = g, acg = (g - ugold)/(gami Delt) + (gami-1)/gami acgold;
        # p: pg = pgold + alfi gami Delt dp; phi, T: a = aold + gami Delt d,
        # ra = (gami-1)/gami raold + d (d = 0 where held)
        du = 1.0 - fu
        cg = (ga - 1.0) / ga
        ca = 1.0 - am / ga
        d_ugold = fu * (gug + Gu) + du * ((1.0 - af) * Gu - am / gdt * Gac - gacg / gdt)
        d_acgold = cg * gacg + ca * Gac
        d_pgold = gpg + Gp
        d_phigold = gphig + Gphi
        d_rphigold = cg * grphig + ca * Grphi
        d_Tgold = gTg + GT
        d_rTgold = cg * grTg + ca * GrT
        d_bc = du * (gug + gacg / gdt + af * Gu + am / gdt * Gac)
        d_amp = float(np.sum(d_bc * self.profile))
        return (d_ugold, d_acgold, d_pgold, d_phigold, d_rphigold, d_Tgold, d_rTgold, Gks,
                d_amp)


class QuenchStep(torch.autograd.Function):
    """(ugold, acgold, pgold, phigold, rphigold, Tgold, rTgold, kscale, amp)
    -> (ug, acg, pg, phig, rphig, Tg, rTg) of one converged coupled step."""

    @staticmethod
    def forward(fc, st, ugold, acgold, pgold, phigold, rphigold, Tgold, rTgold, ks, amp):
        old = [_np(v) for v in (ugold, acgold, pgold, phigold, rphigold, Tgold, rTgold)]
        new = st.solve(old, _np(ks), float(amp))
        out = tuple(_t(v, ugold) for v in new)
        fc.st = st
        fc.save_for_backward(ugold, acgold, pgold, phigold, rphigold, Tgold, rTgold, ks, amp,
                             *out)
        return out

    @staticmethod
    def backward(fc, *gouts):
        st = fc.st
        saved = fc.saved_tensors
        old = [_np(v) for v in saved[:7]]
        ks, amp = _np(saved[7]), float(saved[8])
        new = [_np(v) for v in saved[9:]]
        st.load(old, ks, amp, new)
        seeds = [(_np(gv) if gv is not None else np.zeros_like(nv))
                 for gv, nv in zip(gouts, new)]
        grads = st.step_vjp(seeds)
        like = saved[0]
        out = [_t(gv, like) for gv in grads[:8]]
        g_amp = torch.as_tensor(grads[8], dtype=saved[8].dtype, device=saved[8].device)
        return (None, *out, g_amp)


def quench_step(st, state, ks, amp):
    """``state`` = (ug, acg, pg, phig, rphig, Tg, rTg) of the old level."""
    return QuenchStep.apply(st, *state, ks, amp)


def _quench_segment(st, n, ks, amp, *state):
    for _ in range(n):
        state = QuenchStep.apply(st, *state, ks, amp)
    return state


def run_quench_steps(st, state, ks, amp, nsteps, checkpoint_every=0):
    """``nsteps`` converged coupled steps from ``state``, checkpointed as
    ``run_steps``."""
    state = tuple(state)
    done = 0
    k = int(checkpoint_every)
    while done < nsteps:
        n = min(k, nsteps - done) if k > 0 else nsteps - done
        if k > 0 and torch.is_grad_enabled():
            state = checkpoint(_quench_segment, st, n, ks, amp, *state, use_reentrant=False)
        else:
            state = _quench_segment(st, n, ks, amp, *state)
        done += n
    return state


def quench_heat_loss(st, state, ks, amp, rho_cp, solid_volume, nsteps, checkpoint_every=0):
    """Heat the solid gives off over ``nsteps`` coupled steps,
    sum_i V_i rho_s c_s (T0_i - T_i)."""
    T0 = state[5]
    out = run_quench_steps(st, state, ks, amp, nsteps, checkpoint_every)
    return torch.sum(rho_cp * solid_volume * (T0 - out[5]))
