"""Time steps of the overset pair coupled in one system, on the GPU against
the oracle: a 24x12x12 background and a rotated 8^3 component, Re = 500
(on the component's width) cross-flow, dt = 0.005, 2 steps of 2 Newton
iterations.  Each iteration fills the component's outer boundary from the
background and the background fringe from the component (setBCs_overset,
setBCs.f90:124-167,
with the hex donor search), then assembles both blocks in one launch
(subdomain flags 1 | 2) and solves the coupled system.  The oracle does the
same with its hex donor search (first index wins), its assembly and a
direct sparse solve.  Every step starts from the physical flags (IBC(:,:) = 0
then setBCs_CFD, solveOverset.f90:33-34), so the physical Dirichlet values
never land on the receivers.  Donors bit-exact at every iteration; the
oracle's systems are solved by a sparse LU with FP64 iterative refinement;
FP64 restarted GMRES(30) to a 1e-15 true residual holds every field to
1e-10; the mixed solver at 1e-8 is reported and held to 1e-5.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import direct_solve
from oversetdevfort_amd import _lib as L
from oversetdevfort_amd import configs as C

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-300)


def _alpha(st, p):
    a = {}
    for k in ("ug", "ugm", "dg"):
        a[k] = st[k + "old"] + p["alfi"] * (st[k] - st[k + "old"])
    for k in ("acg", "acgm"):
        a[k] = st[k + "old"] + p["almi"] * (st[k] - st[k + "old"])
    a["pg"] = st["pg"]
    return a


def test_overset_time_steps_match_oracle(oracle_mod):
    from oversetdevfort_amd import solver as S
    O = oracle_mod
    prob = C.overset(nb=(24, 12, 12), nc=8, seed=3)
    m, p = prob.mesh, prob.params
    N = m.nnode
    ncomp = len(prob.fringe["component"][0])
    nback = len(prob.fringe["background"][0])
    print(f"overset pair: {m.nelem} elements, {N} nodes, receivers {ncomp} component, "
          f"{nback} background, theta {np.rad2deg(prob.theta):.1f} deg", flush=True)
    assert ncomp > 0 and nback > 0
    col, row = O.sparsity(N, m.ien)
    prm = O.params_array(p)
    ga, dt = p["gami"], p["delt"]
    st = {k: v.copy() for k, v in prob.fields.items()}
    zeros = np.zeros((N, 3))
    cases = {"fp64 1e-15": (L.OFD_FP64, 1e-15), "mixed 1e-8": (L.OFD_FP32, 1e-8)}
    ctxs = {name: S.setup(prob) for name in cases}

    def consistent_acg(k):
        return (((st["ug"] - st["ugold"]) * (1.0 / dt) + (ga - 1.0) * st["acgold"]) / ga)[k]

    for step in range(2):
        # IBC(:,:) = 0 and the physical flags (setBCs_overset starts with
        # setBCs_CFD): the last step's receivers keep their interpolated
        # velocity through the predictor into the first donor search
        ibc = prob.ibc.copy()
        st["ug"] = st["ugold"].copy()
        st["acg"] = (ga - 1.0) / ga * st["acgold"]
        st["pg"] = st["pgold"].copy()
        mask = ibc[:, :3] == 1
        st["ug"][mask] = prob.bcval[mask]
        st["acg"][mask] = consistent_acg(mask)
        rec = np.concatenate([nodes for nodes, _ in prob.fringe.values()])
        for c in ctxs.values():
            c.reset_ibc()
            S.begin_step(c)
            # the predictor left the receivers at the last step's velocity
            assert np.array_equal(c.get_field("UG")[rec], c.get_field("UGOLD")[rec])
        for it in range(2):
            want = {}
            for name, (nodes, src) in prob.fringe.items():
                # the oracle's donors are 0-based (-1: none), the library's 1-based
                donor, val = O.interp_hex(m, src, zeros, st["ug"], m.xg[nodes], zeros[nodes])
                assert np.all(donor >= 0)
                ibc[nodes, :3] = 1
                st["ug"][nodes] = val
                mask = np.zeros((N, 3), bool)
                mask[nodes] = True
                st["acg"][mask] = consistent_acg(mask)
                want[name] = donor
            R = O.assemble_ns(m, ibc, _alpha(st, p), prm, col, row, 6, 3)
            blocks = np.zeros((row.size, 4, 4))
            blocks[:, :3, :3] = R["k11"].reshape(-1, 3, 3)
            blocks[:, :3, 3] = R["g"]
            blocks[:, 3, :3] = R["d1"]
            blocks[:, 3, 3] = R["m"]
            A = sp.bsr_matrix((blocks, row, col), shape=(4 * N, 4 * N)).tocsc()
            b = np.concatenate([R["rhsu"], R["rhsp"][:, None]], axis=1).ravel()
            x = direct_solve(A, b).reshape(N, 4)
            st["acg"] = st["acg"] + x[:, :3]
            st["ug"] = st["ug"] + ga * dt * x[:, :3]
            st["pg"] = st["pg"] + p["alfi"] * ga * dt * x[:, 3]
            for name, c in ctxs.items():
                prec, rtol = cases[name]
                donors = S.fill_fringe(c, prob.fringe, m.xg)
                for k in want:
                    assert np.array_equal(donors[k] - 1, want[k]), (name, step, it, k)
                info = S.newton_step(c, rtol=rtol, precision=prec, ortho=L.OFD_ORTHO_CGS1,
                                     kspace=30, max_cycles=600, subdomain_flag=3,
                                     raise_noconv=False)
                assert info["relres"] <= max(rtol, 1e-13), (name, step, it, info["relres"])
        for k in ("ug", "acg", "pg"):
            st[k + "old"] = st[k].copy()
        for c in ctxs.values():
            c.advance()
        errs = {name: {k: _rel(c.get_field(k.upper()), st[k]) for k in ("ug", "acg", "pg")}
                for name, c in ctxs.items()}
        print(f"overset step {step + 1} vs oracle: " + "; ".join(
            f"{name}: " + ", ".join(f"{k} {v:.1e}" for k, v in e.items())
            for name, e in errs.items()), flush=True)
        assert max(errs["fp64 1e-15"].values()) <= 1e-10, errs["fp64 1e-15"]
        assert max(errs["mixed 1e-8"].values()) <= 1e-5, errs["mixed 1e-8"]
    for c in ctxs.values():
        c.close()
