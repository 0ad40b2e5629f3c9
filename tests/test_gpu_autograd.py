"""The adjoint behind torch.autograd.Function (no reference code).

* NSResidual: torch.autograd.gradcheck (fast mode, eps 1e-6) and central
  differences on 16 seeded directions per input.
* NavStoStep: gradients of an objective of one converged time step with
  respect to the old state, the viscosity scale and the Dirichlet values,
  against central differences of re-solved steps (eps 1e-6, 16 directions).
* run_steps with a checkpoint every k steps: the gradient equals the one of
  the uncheckpointed chain bit for bit (the kernels are deterministic), each
  segment is re-solved once, and the gradient with respect to the viscosity
  scale and the lid amplitude matches central differences of the whole
  transient.
* TemStep (quench temperature with the flow held): gradients with respect to
  the old temperature and rate, the conductivity scale and the velocity
  against central differences of re-solved steps; run_tem_steps with a
  checkpoint every 3 steps as above, for the conductivity scale and the
  velocity.
"""
import numpy as np
import pytest
import torch

from oversetdevfort_amd import configs as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def A():
    from oversetdevfort_amd import adjoint
    return adjoint


@pytest.fixture(scope="module")
def S():
    from oversetdevfort_amd import solver
    return solver


def _dev():
    return torch.device("cuda:0") if torch.cuda.is_available() else torch.device("cpu")


def _t(a):
    return torch.tensor(np.asarray(a, np.float64), device=_dev())


def _cavity(S, A, n, seed):
    rng = np.random.default_rng(seed)
    prob = C.cavity(n, seed=seed, jitter=0.05)
    N = prob.mesh.nnode
    free = (prob.ibc[:, :3] == 0)
    f = prob.fields
    f["ugold"] = 0.2 * rng.normal(size=(N, 3)) * free + prob.bcval * ~free
    f["acgold"] = 0.5 * rng.normal(size=(N, 3))
    f["pgold"] = rng.normal(size=N)
    f["ug"] = f["ugold"] + 0.05 * rng.normal(size=(N, 3))
    f["acg"] = f["acgold"] + 0.2 * rng.normal(size=(N, 3))
    f["pg"] = f["pgold"] + 0.1 * rng.normal(size=N)
    ctx = S.setup(prob)
    st = A.NSStepper(ctx, prob.ibc, prob.params)
    return st, prob, rng


def _dirs(shape, seed, k=16):
    return [np.random.default_rng(seed + d).normal(size=shape) for d in range(k)]


def _fd_check(fun, x0, grad, dirs, eps=1e-6):
    """Worst relative error of the adjoint directional derivative against
    central differences.  The scale is the larger derivative, floored at 1%
    of |grad| |v|, so that a direction that happens to be almost orthogonal
    to the gradient is judged by the gradient's size and not by its own
    (the difference quotient's rounding does not shrink with it)."""
    worst = 0.0
    gn = float(np.linalg.norm(grad))
    for v in dirs:
        fd = (fun(x0 + eps * v) - fun(x0 - eps * v)) / (2 * eps)
        ad = float(np.sum(grad * v))
        scale = max(abs(fd), abs(ad), 1e-2 * gn * float(np.linalg.norm(v)), 1e-300)
        worst = max(worst, abs(fd - ad) / scale)
    return worst


def test_ns_residual_gradcheck(S, A):
    st, prob, rng = _cavity(S, A, 3, seed=11)
    N = prob.mesh.nnode
    f = prob.fields
    mus = 1.0 + 0.3 * rng.uniform(size=N)
    xs = [_t(f["ug"]).requires_grad_(), _t(f["acg"]).requires_grad_(),
          _t(f["pg"]).requires_grad_(), _t(mus).requires_grad_()]
    ok = torch.autograd.gradcheck(lambda *x: A.ns_residual(st, *x), xs, eps=1e-6,
                                  atol=1e-6, rtol=1e-5, fast_mode=True)
    assert ok
    st.ctx.close()


def test_ns_residual_directional(S, A):
    st, prob, rng = _cavity(S, A, 5, seed=12)
    N = prob.mesh.nnode
    f = prob.fields
    mus = 1.0 + 0.3 * rng.uniform(size=N)
    wu, wp = rng.normal(size=(N, 3)), rng.normal(size=N)
    base = [f["ug"], f["acg"], f["pg"], mus]
    xs = [_t(b).requires_grad_() for b in base]
    ru, rp = A.ns_residual(st, *xs)
    (torch.sum(ru * _t(wu)) + torch.sum(rp * _t(wp))).backward()
    for i, name in enumerate(("ug", "acg", "pg", "mus")):
        def j(x, i=i):
            args = [_t(b) for b in base]
            args[i] = _t(x)
            with torch.no_grad():
                a, b = A.ns_residual(st, *args)
            return float(np.sum(a.cpu().numpy() * wu) + np.sum(b.cpu().numpy() * wp))
        worst = _fd_check(j, base[i], xs[i].grad.cpu().numpy(),
                          _dirs(base[i].shape, 500 + 20 * i))
        print(f"NSResidual d/d{name}: worst relative FD error over 16 directions "
              f"{worst:.2e}", flush=True)
        assert worst <= 1e-6, name
    st.ctx.close()


def _objective(prob, wa, wp):
    free = torch.tensor((prob.ibc[:, :3] == 0).astype(np.float64), device=_dev())
    wa, wp = _t(wa), _t(wp)

    def j(ug, acg, pg):
        return 0.5 * torch.sum((ug * free) ** 2) + torch.sum(wa * acg) + torch.sum(wp * pg)
    return j


def test_nav_sto_step_gradients(S, A):
    st, prob, rng = _cavity(S, A, 5, seed=13)
    N = prob.mesh.nnode
    f = prob.fields
    mus = 1.0 + 0.2 * rng.uniform(size=N)
    J = _objective(prob, 0.1 * rng.normal(size=(N, 3)), 0.01 * rng.normal(size=N))
    base = [f["ugold"], f["acgold"], f["pgold"], mus, prob.bcval]
    xs = [_t(b).requires_grad_() for b in base]
    J(*A.nav_sto_step(st, *xs)).backward()
    dirich = (prob.ibc[:, :3] == 1).astype(np.float64)
    pin = (prob.ibc[:, 3] == 1).astype(np.float64)
    masks = [None, None, pin, None, dirich]
    for i, name in enumerate(("ugold", "acgold", "pgold", "mus", "bcval")):
        def j(x, i=i):
            args = [_t(b) for b in base]
            args[i] = _t(x)
            with torch.no_grad():
                return float(J(*A.nav_sto_step(st, *args)))
        dirs = _dirs(base[i].shape, 600 + 20 * i)
        if masks[i] is not None:
            dirs = [d * masks[i] for d in dirs]
        g = xs[i].grad.cpu().numpy()
        worst = _fd_check(j, base[i], g, dirs)
        print(f"NavStoStep d/d{name}: worst relative FD error over 16 directions "
              f"{worst:.2e}", flush=True)
        assert worst <= 1e-5, name
        if masks[i] is not None:                 # nothing reaches the unused entries
            assert not np.any(g * (1.0 - masks[i]))
    st.ctx.close()


def test_transient_checkpointed(S, A):
    st, prob, rng = _cavity(S, A, 5, seed=14)
    N = prob.mesh.nnode
    f = prob.fields
    mus0 = 1.0 + 0.2 * rng.uniform(size=N)
    J = _objective(prob, 0.05 * rng.normal(size=(N, 3)), 0.0 * rng.normal(size=N))
    nsteps, every = 6, 2
    u0, a0, p0 = _t(f["ugold"]), _t(f["acgold"]), _t(f["pgold"])

    def run(mus, amp, k):
        return A.run_steps(st, u0, a0, p0, mus, amp * _t(prob.bcval), nsteps,
                           checkpoint_every=k)

    grads = {}
    for k in (0, every):
        mus = _t(mus0).requires_grad_()
        amp = torch.tensor(1.0, dtype=torch.float64, device=_dev(), requires_grad=True)
        st.newton_history.clear()
        out = run(mus, amp, k)
        n_fwd = len(st.newton_history)
        J(*out).backward()
        grads[k] = (mus.grad.cpu().numpy(), float(amp.grad))
        solves = len(st.newton_history)
        print(f"checkpoint every {k}: {n_fwd} forward steps, {solves} steps solved in all",
              flush=True)
        assert n_fwd == nsteps
        assert solves == (2 * nsteps if k else nsteps)
    assert np.array_equal(grads[0][0], grads[every][0])
    assert grads[0][1] == grads[every][1]

    def j_of(mus, amp=1.0):
        with torch.no_grad():
            return float(J(*run(_t(mus), amp, 0)))

    worst = _fd_check(j_of, mus0, grads[every][0], _dirs((N,), 700))
    print(f"transient ({nsteps} steps) dj/d(viscosity scale): worst relative FD error over "
          f"16 directions {worst:.2e}", flush=True)
    assert worst <= 1e-5
    eps = 1e-6
    fd = (j_of(mus0, 1.0 + eps) - j_of(mus0, 1.0 - eps)) / (2 * eps)
    print(f"transient dj/d(lid amplitude): adjoint {grads[every][1]:.10e} FD {fd:.10e}",
          flush=True)
    assert abs(fd - grads[every][1]) <= 1e-5 * abs(fd)
    st.ctx.close()


def _tank(S, A, n, seed):
    """Quench tank with the flow held: a seeded velocity and VOF field, old
    temperatures on both sides of T_sat, T Dirichlet nodes on the floor.
    Every material property (T_sat included) is of order one and the step
    is 10 ms, so that conduction moves the temperature within a step by more
    than central differences of re-solved steps can resolve (with water's
    rho cp of 4.2e6 a 1 ms step changes T by ~1e-10 per unit conductivity
    scale, and at T ~ 400 K one unit in the last place of T is 6e-14)."""
    rng = np.random.default_rng(seed)
    prob = C.quench(n, seed=seed, jitter=0.05)
    prob.material.update(rhoa=0.5, rhow=1.0, cpa=0.8, cpw=1.0, kappaa=0.05, kappaw=0.2,
                         rhos=2.0, cps=0.5, kappas=1.0, lh=20.0, Ts=1.0, tref=1.0)
    prob.params["delt"] = 1e-2
    N = prob.mesh.nnode
    x = prob.mesh.xg
    prob.ibc[:, 5] = (x[:, 2] < 1e-9).astype(np.int32)
    phi = np.clip(prob.fields["phig"] + 0.1 * rng.uniform(size=N), 0.0, 1.0)
    ctx = S.setup(prob)
    ctx.set_quench_params(prob.material)
    st = A.TemStepper(ctx, prob.ibc, prob.params, phi)
    T0 = rng.uniform(0.9, 1.1, size=N)         # both sides of T_sat = 1
    rT0 = 0.03 * rng.normal(size=N)
    u = 0.05 * rng.normal(size=(N, 3))
    ks = 1.0 + 0.2 * rng.uniform(size=N)
    return st, prob, rng, T0, rT0, u, ks


def test_tem_step_gradients(S, A):
    st, prob, rng, T0, rT0, u, ks = _tank(S, A, 5, seed=31)
    N = prob.mesh.nnode
    w, wr = _t(rng.normal(size=N)), _t(1e-2 * rng.normal(size=N))

    def J(T, rT):
        return torch.sum(w * T) + torch.sum((T - 1.0) ** 2) + torch.sum(wr * rT)

    base = [T0, rT0, ks, u]
    xs = [_t(b).requires_grad_() for b in base]
    J(*A.tem_step(st, *xs)).backward()
    print("temperature Newton residuals:", ["%.1e" % h for h in st.newton_history[-1]],
          flush=True)
    for i, name in enumerate(("Tgold", "rTgold", "kscale", "u")):
        def j(x, i=i):
            args = [_t(b) for b in base]
            args[i] = _t(x)
            with torch.no_grad():
                return float(J(*A.tem_step(st, *args)))
        worst = _fd_check(j, base[i], xs[i].grad.cpu().numpy(),
                          _dirs(base[i].shape, 1000 + 20 * i))
        print(f"TemStep d/d{name}: worst relative FD error over 16 directions {worst:.2e}",
              flush=True)
        assert worst <= 1e-5, name
    st.ctx.close()


def test_tem_transient_checkpointed(S, A):
    st, prob, rng, T0, rT0, u, ks0 = _tank(S, A, 5, seed=32)
    N = prob.mesh.nnode
    w = _t(rng.normal(size=N))
    nsteps, every = 6, 3

    def J(T, rT):
        return torch.sum(w * T) + torch.sum((T - 1.0) ** 2)

    def run(ks, uu, k):
        return A.run_tem_steps(st, _t(T0), _t(rT0), ks, uu, nsteps, checkpoint_every=k)

    grads = {}
    for k in (0, every):
        ks = _t(ks0).requires_grad_()
        uu = _t(u).requires_grad_()
        st.newton_history.clear()
        out = run(ks, uu, k)
        n_fwd = len(st.newton_history)
        J(*out).backward()
        grads[k] = (ks.grad.cpu().numpy(), uu.grad.cpu().numpy())
        solves = len(st.newton_history)
        print(f"temperature, checkpoint every {k}: {n_fwd} forward steps, {solves} steps "
              f"solved in all", flush=True)
        assert n_fwd == nsteps
        assert solves == (2 * nsteps if k else nsteps)
    assert np.array_equal(grads[0][0], grads[every][0])
    assert np.array_equal(grads[0][1], grads[every][1])

    def j_of(ks, uu=u):
        with torch.no_grad():
            return float(J(*run(_t(ks), _t(uu), 0)))

    worst = _fd_check(j_of, ks0, grads[every][0], _dirs((N,), 1100))
    print(f"temperature transient ({nsteps} steps) dj/d(conductivity scale): worst relative "
          f"FD error over 16 directions {worst:.2e}", flush=True)
    assert worst <= 1e-5
    worst = _fd_check(lambda uu: j_of(ks0, uu), u, grads[every][1], _dirs((N, 3), 1200, 4))
    print(f"temperature transient dj/du: worst relative FD error over 4 directions "
          f"{worst:.2e}", flush=True)
    assert worst <= 1e-5
    st.ctx.close()
