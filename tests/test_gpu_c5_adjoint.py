"""The adjoint configuration on the GPU (no reference code): the quench tank
on 128x64x64 hexes (524,288 elements), 20 temperature steps at dt = 1e-3
with the flow held and a checkpoint every 5 steps.  The objective is the
heat the solid block gives off, sum_i V_i rho_s c_s (T0_i - T_i) over its
lumped nodal volumes, i.e. by the solid's energy balance the
time-integrated heat flux across its surface.  Its gradients with respect
to the per-node conductivity scale (exp(0.1 z), z seeded) and the amplitude
of the held velocity are checked against central differences of whole
re-run transients (eps 1e-6, 16 seeded directions), with FP64 adjoint
solves and with FP32-matrix inner solves.

The tolerance is 1e-4 on the error relative to the larger of the two
derivatives (floored at 1% of |grad| |v|): with T up to 900 K one unit in
the last place of T is 1.1e-13, and the heat sum over ~5,000 solid nodes
carries ~1e-10 J of rounding against differences of 2e-6 in the scale.
"""
import numpy as np
import pytest
import torch

from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C

pytestmark = pytest.mark.gpu


def test_c5_heat_loss_gradient():
    from oversetdevfort_amd import adjoint as A
    from oversetdevfort_amd import solver as S
    prob = C.adjoint_quench()
    N = prob.mesh.nnode
    print(f"adjoint tank: {prob.mesh.nelem} elements, {N} nodes", flush=True)
    ctx = S.setup(prob)
    ctx.set_quench_params(prob.material)
    dev = torch.device("cuda:0") if torch.cuda.is_available() else torch.device("cpu")

    def t(a):
        return torch.tensor(np.asarray(a, np.float64), device=dev)

    m = prob.material
    rc = m["rhos"] * m["cps"]
    T0, rT0 = t(prob.fields["Tgold"]), t(np.zeros(N))
    circ, vol = t(prob.circulation), t(prob.solid_volume)
    nsteps, every = 20, 5
    grads = {}
    for tag, aprec in (("fp64", L.OFD_FP64), ("fp32", L.OFD_FP32)):
        st = A.TemStepper(ctx, prob.ibc, prob.params, prob.fields["phig"],
                          adj_precision=aprec)
        ks = t(prob.kscale).requires_grad_()
        amp = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
        J = A.solid_heat_loss(st, T0, rT0, ks, amp * circ, rc, vol, nsteps, every)
        J.backward()
        grads[tag] = (ks.grad.cpu().numpy(), float(amp.grad), float(J.detach()))
        assert len(st.newton_history) == 2 * nsteps
        print(f"{tag} adjoint: heat given off {float(J.detach()):.6e} J, dJ/d(amplitude) "
              f"{float(amp.grad):.8e}, |dJ/d(scale)| {np.linalg.norm(grads[tag][0]):.6e}",
              flush=True)
    d32 = (np.linalg.norm(grads["fp32"][0] - grads["fp64"][0]) /
           np.linalg.norm(grads["fp64"][0]))
    print(f"FP32-matrix inner solves vs FP64: gradient differs by {d32:.2e}", flush=True)
    assert d32 <= 1e-8

    st = A.TemStepper(ctx, prob.ibc, prob.params, prob.fields["phig"])

    def j_of(ks, amp=1.0):
        with torch.no_grad():
            return float(A.solid_heat_loss(st, T0, rT0, t(ks), amp * circ, rc, vol, nsteps))

    eps = 1e-6
    g, gn = grads["fp64"][0], np.linalg.norm(grads["fp64"][0])
    worst = {"fp64": 0.0, "fp32": 0.0}
    for d in range(16):
        v = np.random.default_rng(1300 + d).standard_normal(N)
        fd = (j_of(prob.kscale + eps * v) - j_of(prob.kscale - eps * v)) / (2 * eps)
        for tag in worst:
            ad = float(grads[tag][0] @ v)
            scale = max(abs(fd), abs(ad), 1e-2 * gn * np.linalg.norm(v))
            worst[tag] = max(worst[tag], abs(fd - ad) / scale)
    print(f"dJ/d(conductivity scale): worst relative error over 16 directions, FP64 adjoint "
          f"{worst['fp64']:.2e}, FP32-matrix {worst['fp32']:.2e}", flush=True)
    fd = (j_of(prob.kscale, 1.0 + eps) - j_of(prob.kscale, 1.0 - eps)) / (2 * eps)
    ea = abs(fd - grads["fp64"][1]) / abs(fd)
    print(f"dJ/d(amplitude): adjoint {grads['fp64'][1]:.8e}, differences {fd:.8e} "
          f"(relative {ea:.1e})", flush=True)
    ctx.close()
    assert worst["fp64"] <= 1e-4 and worst["fp32"] <= 1e-4
    assert ea <= 1e-4
