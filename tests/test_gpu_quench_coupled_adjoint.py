"""The coupled quench adjoint on the GPU (no reference counterpart): flow,
VOF field and temperature differentiated together through converged time
steps of the staggered iteration (solveFlow.f90:57-161) with exact Newton
steps, on the quench tank with an inlet (configs.quench_inflow: inflow
amplitude times a sine bump on a patch of the wall x = 0, open top, solid
block at 900 K), a diffuse interface and phase-change coefficients 1e-3
(with the tank's 0.1 the exact Newton iteration of a step diverges, so its
root, which the implicit-function-theorem adjoint needs, is not found).

The tank of 8^3 trilinear hexes and the same tank with each hex split into
6 linear tets or 2 linear prisms.

Checks, against central differences of re-solved steps (the forward is
converged to its rounding floor, about 1e-12 of the predictor's residual;
steps of 1e-6, larger along directions whose derivative is small against
the objective):
* one step, a random linear functional of every new field, with respect to
  the inflow amplitude and each old field along random directions, and a
  random functional of the temperature change with respect to the
  conductivity scale (the full functional's derivative along the scale is
  of the order of its convergence floor on the tets);
* four steps of the heat the solid gives off, with a checkpoint every 2
  steps, with respect to the amplitude and the conductivity scale; the
  checkpointed and plain reverse sweeps give the same gradient bit for bit.
"""
import numpy as np
import pytest
import torch

from oversetdevfort_amd import adjoint as A
from oversetdevfort_amd import configs as C

pytestmark = pytest.mark.gpu

NAMES = ("ug", "acg", "pg", "phig", "rphig", "Tg", "rTg")


@pytest.fixture(scope="module", params=["hex", "tet", "prism"])
def setup(request):
    from oversetdevfort_amd import solver as S
    prob = C.quench_inflow((8, 8, 8), seed=5, amp=0.05, diffuse=0.08, rate=1e-3,
                           elem=request.param)
    ctx = S.setup(prob)
    ctx.set_quench_params(prob.material)
    st = A.QuenchStepper(ctx, prob.ibc, prob.params, prob.inflow_profile)
    yield prob, ctx, st
    ctx.close()


def _state(prob):
    f = prob.fields
    return tuple(torch.tensor(f[k + "old"] if k + "old" in f else f[k], dtype=torch.float64)
                 for k in NAMES)


def _fd(fun, x0, dx, eps):
    return (fun(x0 + eps * dx) - fun(x0 - eps * dx)) / (2.0 * eps)


def _step(J, ad, lo=1e-6, hi=1e-2):
    """Difference step along a direction: at least 1e-6, larger where the
    directional derivative is small against the objective, so that the
    change (about 1e-5 of |J|) stands above the objective's rounding floor
    (the converged steps hold it to about 1e-12)."""
    return float(np.clip(1e-5 * abs(J) / max(abs(ad), 1e-300), lo, hi))


def test_one_step_gradients(setup):
    prob, ctx, st = setup
    rng = np.random.default_rng(11)
    state = _state(prob)
    ks0 = torch.tensor(prob.kscale)
    amp0 = torch.tensor(prob.amp, dtype=torch.float64)
    scale = {"ug": 0.05, "acg": 1.0, "pg": 1e3, "phig": 0.1, "rphig": 1.0, "Tg": 100.0,
             "rTg": 1e3}
    W = [torch.tensor(rng.normal(size=s.shape)) / scale[n] for s, n in zip(state, NAMES)]

    def objective(old, ks, amp):
        new = A.quench_step(st, old, ks, amp)
        return sum(torch.sum(w * v) for w, v in zip(W, new))

    old = tuple(s.clone().requires_grad_(True) for s in state)
    ks = ks0.clone().requires_grad_(True)
    amp = amp0.clone().requires_grad_(True)
    J = objective(old, ks, amp)
    J.backward()
    hist = st.newton_history[-1]
    print(f"one step: J = {J.item():.10e}; Newton residuals (u, p, phi, T) "
          f"{hist[0]} -> {hist[-1]} in {len(hist) - 1} iterations; adjoint sweeps "
          f"{st.gs_history[-1]}", flush=True)
    # amplitude
    f_amp = lambda a: objective(state, ks0, a).item()
    fd = _fd(f_amp, amp0, torch.tensor(1.0, dtype=torch.float64), 1e-6)
    e = abs(fd - amp.grad.item()) / abs(fd)
    print(f"dJ/damp: adjoint {amp.grad.item():.10e} differences {fd:.10e} rel {e:.2e}",
          flush=True)
    assert e <= 1e-4
    # conductivity scale: the functional's derivative is small against its
    # convergence floor (about 1e-12 of each field), so it is checked on the
    # temperature change, which the scale moves and the pressure and velocity
    # terms do not swamp
    v = torch.tensor(rng.normal(size=ks0.shape))
    iT = NAMES.index("Tg")

    def objective_T(ks):
        new = A.quench_step(st, state, ks, amp0)
        return torch.sum(W[iT] * (new[iT] - state[iT]))

    ksT = ks0.clone().requires_grad_(True)
    JT = objective_T(ksT)
    JT.backward()
    ad = torch.sum(ksT.grad * v).item()
    eps = _step(JT.item(), ad)
    fd = _fd(lambda k: objective_T(k).item(), ks0, v, eps)
    e = abs(fd - ad) / abs(fd)
    print(f"dJ_T/dks . v: adjoint {ad:.10e} differences {fd:.10e} (step {eps:.1e}) "
          f"rel {e:.2e}; the full functional's dJ/dks . v {torch.sum(ks.grad * v).item():.3e}",
          flush=True)
    assert e <= 1e-4
    # each old field
    for i, n in enumerate(NAMES):
        v = torch.tensor(rng.normal(size=state[i].shape)) * scale[n]

        def f_old(x, i=i):
            s = list(state)
            s[i] = x
            return objective(tuple(s), ks0, amp0).item()

        ad = torch.sum(old[i].grad * v).item()
        eps = _step(J.item(), ad)
        fd = _fd(f_old, state[i], v, eps)
        e = abs(fd - ad) / max(abs(fd), 1e-300)
        print(f"dJ/d{n}old . v: adjoint {ad:.10e} differences {fd:.10e} (step {eps:.1e}) "
              f"rel {e:.2e}", flush=True)
        assert e <= 1e-4


def test_transient_heat_loss_gradient(setup):
    prob, ctx, st = setup
    rng = np.random.default_rng(12)
    state = _state(prob)
    ks0 = torch.tensor(prob.kscale)
    amp0 = torch.tensor(prob.amp, dtype=torch.float64)
    m = prob.material
    rho_cp = m["rhos"] * m["cps"]
    vol = torch.tensor(prob.solid_volume)
    nsteps = 4

    grads = []
    for ck in (2, 0):
        ks = ks0.clone().requires_grad_(True)
        amp = amp0.clone().requires_grad_(True)
        Q = A.quench_heat_loss(st, state, ks, amp, rho_cp, vol, nsteps, checkpoint_every=ck)
        Q.backward()
        grads.append((ks.grad.clone(), amp.grad.clone()))
        print(f"heat loss over {nsteps} steps (checkpoint every {ck}): {Q.item():.10e} J; "
              f"dQ/damp {amp.grad.item():.10e}", flush=True)
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])
    ks_g, amp_g = grads[0]

    def Qf(ks, amp):
        with torch.no_grad():
            return A.quench_heat_loss(st, state, ks, amp, rho_cp, vol, nsteps).item()

    fd = _fd(lambda a: Qf(ks0, a), amp0, torch.tensor(1.0, dtype=torch.float64), 1e-6)
    e = abs(fd - amp_g.item()) / abs(fd)
    print(f"dQ/damp: adjoint {amp_g.item():.10e} differences {fd:.10e} rel {e:.2e}", flush=True)
    assert e <= 1e-4
    Q0 = Qf(ks0, amp0)
    for t in range(2):
        v = torch.tensor(rng.normal(size=ks0.shape))
        ad = torch.sum(ks_g * v).item()
        eps = _step(Q0, ad)
        fd = _fd(lambda k: Qf(k, amp0), ks0, v, eps)
        e = abs(fd - ad) / abs(fd)
        print(f"dQ/dks . v[{t}]: adjoint {ad:.10e} differences {fd:.10e} (step {eps:.1e}) "
              f"rel {e:.2e}", flush=True)
        assert e <= 1e-4
