"""Synthetic problem set-ups: meshes, Dirichlet flags, initial state, parameters.

Host-side set-up that the reference does through param.dat, the mesh files
and setBCs_CFD; the GPU path consumes the resulting arrays.

* ``cavity``  -- 3D lid-driven cavity, unit cube, n^3 trilinear hexes with
  interior nodes jittered by U(-5%, 5%) of h; Re = 100 (rho 1, mu 0.01), lid
  velocity (1,0,0) on z = 1, no-slip elsewhere, at rest; dt = 0.01.
* ``channel`` -- [0,4]x[0,1]x[0,1] on nx*ny*nz hexes; Poiseuille inflow of
  peak 1 at x = 0, no-slip walls y,z in {0,1}, natural outflow at x = 4;
  initial field = Poiseuille profile + a seeded divergence-free sinusoidal
  perturbation of amplitude 0.01; Re = 1000 (rho 1, mu 1e-3); dt = 0.005.

* ``quench`` -- n^3 hex tank on the unit cube: water below z = 0.6 with a
  seeded vapour sphere of radius 0.1, vapour above; density ratio 1000,
  viscosity ratio 100; a solid block [0.4,0.6]^3 (ELM_ID 101, fluid 102 as
  solveQuenching.f90:693) at 900 K in liquid at 300 K; piecewise-constant
  k and c_p; dt = 1e-3.  The reference's kernels treat the VOF field phi as
  the gas fraction (prop_interp(rhow, rhoa, phi), evaporation ~ (1 - phi)),
  so the liquid is phi = 0.

* ``quench_inflow`` -- the quench tank with an inlet patch on the wall x = 0
  (below the liquid level) whose normal velocity is an amplitude times a
  sine bump; the fluid leaves through the open top.  The coupled quench
  adjoint's configuration (gradients with respect to the inflow amplitude
  and a per-node conductivity scale).

* ``overset`` -- background [0,2]x[0,1]x[0,1] on nb hexes (ELM_ID 1) plus a
  cube component of side 0.4 on nc^3 hexes (ELM_ID 2), rotated by a seeded
  angle in [0, 30] degrees about z and centred at (0.9, 0.5, 0.5), in one
  mesh of two disconnected blocks.  Re = 500 cross-flow (u = (1, 0, 0) at
  x = 0, slip walls y, z, natural outflow at x = 2; rho 1, mu 8e-4 on the
  component side 0.4); dt = 0.005.  Fringe receivers: the component's
  outer-boundary nodes (donors in the background, interpolation.f90:46-60)
  and, as the background-side fringe the reference leaves to pre-cut
  meshes, the background nodes inside the component at a depth between two
  component cells and two component cells plus one background cell (donors
  in the component).

All use generalized-alpha with rho_inf = 0.5 (param.f90:469-480) and the
reference's hard-coded gravity (0, 0, -9.81) (param.f90:253-257).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import mesh as M


def genalpha(rhoinf):
    """param.f90:469-480."""
    if rhoinf < 0.0 or rhoinf > 1.0:
        return dict(almi=1.0, alfi=1.0, gami=1.0, beti=1.0)
    almi = (3.0 - rhoinf) / (1.0 + rhoinf) / 2.0
    alfi = 1.0 / (1.0 + rhoinf)
    gami = 0.5 + almi - alfi
    beti = 0.25 * (1.0 + almi - alfi) * (1.0 + almi - alfi)
    return dict(almi=almi, alfi=alfi, gami=gami, beti=beti)


@dataclasses.dataclass
class Problem:
    name: str
    mesh: M.Mesh
    ibc: np.ndarray        # (N, 4) int32, columns u, v, w, p
    bcval: np.ndarray      # (N, 3) Dirichlet velocity values
    fields: dict           # initial node fields, reference names (lower case)
    params: dict

    @property
    def dirichlet_mask(self):
        return self.ibc[:, :3] == 1


def _zero_fields(N):
    z3 = np.zeros((N, 3))
    return {k: z3.copy() for k in ("dg", "dgold", "ug", "ugold", "ugm", "ugmold",
                                   "acg", "acgold", "acgm", "acgmold")} | {
        "pg": np.zeros(N), "pgold": np.zeros(N)}


def cavity(n=16, seed=1, rhoinf=0.5, jitter=0.05) -> Problem:
    mesh = M.structured_hex(n, n, n, jitter=jitter, seed=seed)
    N = mesh.nnode
    faces = M.boundary_masks(mesh, (0, 0, 0), (1, 1, 1))
    wall = np.zeros(N, bool)
    for v in faces.values():
        wall |= v
    ibc = np.zeros((N, 4), np.int32)
    ibc[wall, :3] = 1
    bcval = np.zeros((N, 3))
    bcval[faces["zhi"], 0] = 1.0      # the lid wins on its edges
    ibc[0, 3] = 1                     # pressure fixed at node 0 (enclosed flow)
    f = _zero_fields(N)
    f["ug"][wall] = bcval[wall]
    f["ugold"][:] = f["ug"]
    p = dict(genalpha(rhoinf), delt=0.01, rho=1.0, mu=0.01, kdc=0.0,
             grav=(0.0, 0.0, -9.81))
    return Problem("cavity", mesh, ibc, bcval, f, p)


def weak_cavity(n=16, seed=1, rhoinf=0.5, jitter=0.05) -> Problem:
    """The lid-driven cavity with the no-slip walls imposed weakly: the lid
    keeps its Dirichlet velocity; the other five walls carry no velocity IBC
    and are weak-Dirichlet boundaries (BCugType 2, FaceAssembly_NS_weak)
    whose wall velocity is the (zero) mesh velocity.  Pressure held at node 0."""
    prob = cavity(n, seed, rhoinf, jitter)
    mesh = prob.mesh
    faces = M.boundary_masks(mesh, (0, 0, 0), (1, 1, 1))
    prob.ibc[:, :3] = 0
    prob.ibc[faces["zhi"], :3] = 1
    prob.weak_faces = M.box_faces(mesh, ("xlo", "xhi", "ylo", "yhi", "zlo"))
    prob.name = "weak_cavity"
    return prob


def from_part_file(path, params, fem_flag=1, weak_face_ids=()) -> Problem:
    """A problem on the mesh of a reference part file (input_mesh,
    input.f90:15-135): zero fields, no Dirichlet flags (setBCs sets them from
    the run's parameters, which the file does not hold), and the faces of the
    boundaries listed in ``weak_face_ids`` registered as weak-Dirichlet
    (BCugType 2)."""
    from . import refio
    r = refio.read_part(path, fem_flag)
    mesh = r["mesh"]
    N = mesh.nnode
    prob = Problem("part", mesh, np.zeros((N, 4), np.int32), np.zeros((N, 3)),
                   _zero_fields(N), dict(params))
    weak = [b for b in r["bounds"] if b["face_id"] in set(weak_face_ids)]
    if weak:
        prob.weak_faces = (np.concatenate([b["f2e"] for b in weak]),
                           np.concatenate([b["face_or"] for b in weak]))
    prob.node_id = r["node_id"]
    prob.bounds = r["bounds"]
    return prob


def weak_channel(nx=64, ny=32, nz=32, seed=1, rhoinf=0.5) -> Problem:
    """The channel with its four walls weak-Dirichlet (BCugType 2,
    FaceAssembly_NS_weak + FaceAssembly_NS_outflow over their faces) instead
    of strong; the inflow keeps its Dirichlet profile."""
    prob = channel(nx, ny, nz, seed, rhoinf)
    mesh = prob.mesh
    b = M.boundary_masks(mesh, (0.0, 0.0, 0.0), (4.0, 1.0, 1.0))
    wall = b["ylo"] | b["yhi"] | b["zlo"] | b["zhi"]
    prob.ibc[wall & ~b["xlo"], :3] = 0
    prob.weak_faces = M.box_faces(mesh, ("ylo", "yhi", "zlo", "zhi"))
    prob.name = "weak_channel"
    return prob


def poiseuille(y, z):
    return 16.0 * y * (1.0 - y) * z * (1.0 - z)


def div_free_perturbation(x, lengths, seed, amp=0.01, nmodes=4, raw=False):
    """u' = curl(psi), psi = sum_m c_m sin(2 pi k_m . x / L + phi_m), scaled so
    that max |u'| over the given points is ``amp`` (unscaled when ``raw``)."""
    rng = np.random.default_rng(seed)
    u = np.zeros_like(x)
    Lv = np.asarray(lengths, dtype=np.float64)
    for _ in range(nmodes):
        k = rng.integers(1, 4, size=3).astype(np.float64)
        c = rng.normal(size=3)
        phi = rng.uniform(0, 2 * np.pi)
        g = 2 * np.pi * k / Lv
        th = x @ g + phi
        u += np.cos(th)[:, None] * np.cross(g, c)[None, :]
    return u if raw else amp * u / np.abs(u).max()


def channel(nx=64, ny=32, nz=32, seed=1, rhoinf=0.5) -> Problem:
    lo, hi = (0.0, 0.0, 0.0), (4.0, 1.0, 1.0)
    mesh = M.structured_hex(nx, ny, nz, lo=lo, hi=hi, jitter=0.0, seed=seed)
    N = mesh.nnode
    x = mesh.xg
    faces = M.boundary_masks(mesh, lo, hi)
    wall = faces["ylo"] | faces["yhi"] | faces["zlo"] | faces["zhi"]
    inflow = faces["xlo"]
    ibc = np.zeros((N, 4), np.int32)
    ibc[wall | inflow, :3] = 1
    bcval = np.zeros((N, 3))
    bcval[inflow, 0] = poiseuille(x[inflow, 1], x[inflow, 2])
    bcval[wall] = 0.0
    f = _zero_fields(N)
    u0 = np.zeros((N, 3))
    u0[:, 0] = poiseuille(x[:, 1], x[:, 2])
    u0 += div_free_perturbation(x, (4.0, 1.0, 1.0), seed)
    d = ibc[:, :3] == 1
    u0[d] = bcval[d]
    f["ug"][:] = u0
    f["ugold"][:] = u0
    p = dict(genalpha(rhoinf), delt=0.005, rho=1.0, mu=1e-3, kdc=0.0,
             grav=(0.0, 0.0, -9.81))
    return Problem("channel", mesh, ibc, bcval, f, p)


def apply_dirichlet(ug, acg, ugold, acgold, prob: Problem):
    """setBCs_CFD value part (setBCs.f90:84-91): Dirichlet velocity and the
    consistent acceleration acg = ((u - uold)/dt + (gami-1) acgold)/gami."""
    p = prob.params
    m = prob.dirichlet_mask
    ug = ug.copy()
    acg = acg.copy()
    ug[m] = prob.bcval[m]
    acg[m] = (((ug - ugold) * (1.0 / p["delt"]) + (p["gami"] - 1.0) * acgold) /
              p["gami"])[m]
    return ug, acg


# water / vapour / solid data for the quenching tank (param.f90:222-242)
QUENCH_MATERIAL = dict(rhoa=1.0, rhow=1000.0, mua=1.0e-5, muw=1.0e-3, cpa=2000.0,
                       cpw=4200.0, kappaa=0.025, kappaw=0.6, rhos=7800.0, mus=0.0,
                       cps=500.0, kappas=30.0, Ts=373.15, lh=2.26e6, c_cond=0.1,
                       c_evap=0.1, c_dmdot=1.0, tem_kdc=0.0, tref=425.0, fluid_id=102,
                       lsc_kdc=0.0)


def quench(n=48, seed=1, rhoinf=0.5, jitter=0.0, dims=None) -> Problem:
    """The quench tank on the unit cube, n^3 hexes (or ``dims``)."""
    nx, ny, nz = dims if dims is not None else (n, n, n)
    mesh = M.structured_hex(nx, ny, nz, jitter=jitter, seed=seed)
    N, E = mesh.nnode, mesh.nelem
    x = mesh.xg
    rng = np.random.default_rng(seed)
    # element ids: solid block [0.4, 0.6]^3 by element centroid
    cen = x[mesh.ien].mean(axis=1)
    solid = np.all((cen > 0.4) & (cen < 0.6), axis=1)
    mesh.elm_id = np.where(solid, 101, 102).astype(np.int32)
    # gas fraction: 0 in the liquid below z = 0.6, 1 above and in a bubble
    c = rng.uniform([0.15, 0.15, 0.15], [0.85, 0.85, 0.45])
    while np.all(np.abs(c - 0.5) < 0.2):      # keep the bubble off the block
        c = rng.uniform([0.15, 0.15, 0.15], [0.85, 0.85, 0.45])
    phi = (x[:, 2] >= 0.6).astype(np.float64)
    phi[np.linalg.norm(x - c, axis=1) < 0.1] = 1.0
    # temperature: 900 K on the nodes of solid elements, 300 K elsewhere
    T = np.full(N, 300.0)
    T[np.unique(mesh.ien[solid].ravel())] = 900.0
    faces = M.boundary_masks(mesh, (0, 0, 0), (1, 1, 1))
    wall = np.zeros(N, bool)
    for k in ("xlo", "xhi", "ylo", "yhi", "zlo"):
        wall |= faces[k]
    ibc = np.zeros((N, 8), np.int32)
    ibc[wall, :3] = 1                # no-slip tank walls; open top
    # nodes inside the solid block: the NS + VOF element skips solid
    # elements, so a node no fluid element touches is held in u, p and phi
    # (BCLhs_NS_conv columns 1-3, 7, 8)
    wet = np.zeros(N, bool)
    wet[mesh.ien[~solid].ravel()] = True
    ibc[np.ix_(~wet, [0, 1, 2, 6, 7])] = 1
    f = _zero_fields(N)
    f.update(phig=phi.copy(), phigold=phi.copy(), rphig=np.zeros(N), rphigold=np.zeros(N),
             Tg=T.copy(), Tgold=T.copy(), rTg=np.zeros(N), rTgold=np.zeros(N))
    rho = QUENCH_MATERIAL["rhow"]
    p = dict(genalpha(rhoinf), delt=1e-3, rho=rho, mu=QUENCH_MATERIAL["muw"], kdc=0.0,
             grav=(0.0, 0.0, -9.81))
    prob = Problem("quench", mesh, ibc, np.zeros((N, 3)), f, p)
    prob.material = dict(QUENCH_MATERIAL)
    return prob


# linhexshl corner signs (xi, eta, zeta) in the reference node order
_HEX_SGN = np.array([[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1],
                     [-1, -1, -1], [1, -1, -1], [1, -1, 1], [-1, -1, 1]], np.float64)


def lumped_volumes(mesh, elements=None):
    """Row sums of the mass matrix, sum_e int N_a dV, over all elements or
    the boolean selection ``elements``: hexes by the 2x2x2 Gauss rule, linear
    prisms by the 6-point rule (linprishl's shape functions), linear tets
    exactly (a quarter of the element's volume at each node)."""
    g = 1.0 / np.sqrt(3.0)
    ien = mesh.ien if elements is None else mesh.ien[elements]
    X = mesh.xg[ien]                                   # (E, nshl, 3)
    vol = np.zeros(mesh.nnode)
    if ien.shape[1] == 4:
        d = np.abs(np.linalg.det(X[:, :3] - X[:, 3:4])) / 6.0
        np.add.at(vol, ien.ravel(), np.repeat(0.25 * d, 4))
        return vol
    if ien.shape[1] == 6:
        for q in range(6):
            zi = 2.0 / 3.0 if q % 3 == 0 else 1.0 / 6.0
            eta = 2.0 / 3.0 if q % 3 == 1 else 1.0 / 6.0
            zeta = -g if q < 3 else g
            tri = np.array([zi, eta, 1.0 - zi - eta])
            dtri = np.array([[1.0, 0.0], [0.0, 1.0], [-1.0, -1.0]])
            lo, hi = 0.5 * (1.0 - zeta), 0.5 * (1.0 + zeta)
            N = np.concatenate([tri * lo, tri * hi])
            dN = np.concatenate([np.column_stack([dtri * lo, -0.5 * tri]),
                                 np.column_stack([dtri * hi, 0.5 * tri])])   # (6, 3)
            J = np.einsum("eai,aj->eij", X, dN)
            dv = np.abs(np.linalg.det(J)) / 6.0
            np.add.at(vol, ien.ravel(), (dv[:, None] * N[None, :]).ravel())
        return vol
    for q in range(8):
        xi = np.array([g if q & 1 else -g, g if q & 2 else -g, g if q & 4 else -g])
        f = 1.0 + _HEX_SGN * xi                        # (8, 3)
        N = 0.125 * f.prod(axis=1)
        dN = np.stack([0.125 * _HEX_SGN[:, 0] * f[:, 1] * f[:, 2],
                       0.125 * _HEX_SGN[:, 1] * f[:, 0] * f[:, 2],
                       0.125 * _HEX_SGN[:, 2] * f[:, 0] * f[:, 1]], axis=1)   # (8, 3)
        J = np.einsum("eai,aj->eij", X, dN)
        dv = np.linalg.det(J)
        np.add.at(vol, ien.ravel(), (dv[:, None] * N[None, :]).ravel())
    return vol


def circulation(x):
    """A divergence-free swirl about the tank's vertical axis, tangential at
    the walls, unit amplitude: u = curl(0, 0, psi),
    psi = sin(pi x) sin(pi y) sin(pi z) / pi."""
    sx, sy, sz = (np.sin(np.pi * x[:, k]) for k in range(3))
    cx, cy = np.cos(np.pi * x[:, 0]), np.cos(np.pi * x[:, 1])
    return np.stack([sx * cy * sz, -cx * sy * sz, np.zeros(len(x))], axis=1)


def adjoint_quench(dims=(128, 64, 64), seed=5, speed=0.05):
    """The adjoint configuration: the quench tank on ``dims`` hexes with the
    flow held (VOF field as set up, the velocity ``speed`` x circulation),
    a per-node conductivity scale exp(0.1 z), z seeded standard normal, and
    the solid block's lumped nodal volumes for the heat it gives off."""
    prob = quench(dims=dims, seed=seed)
    rng = np.random.default_rng(seed)
    N = prob.mesh.nnode
    prob.kscale = np.exp(0.1 * rng.standard_normal(N))
    prob.circulation = speed * circulation(prob.mesh.xg)
    prob.solid_volume = lumped_volumes(prob.mesh, prob.mesh.elm_id == 101)
    return prob


def inflow_bump(x, lo=(0.3, 0.1), hi=(0.7, 0.4)):
    """Profile of the tank's inlet: on the face x = 0, the patch y in
    [lo0, hi0], z in [lo1, hi1] (below the liquid level), the normal velocity
    sin(pi (y - lo0) / (hi0 - lo0)) sin(pi (z - lo1) / (hi1 - lo1)), zero on
    the patch's edges and outside it; (N, 3) with the x component only."""
    y, z = x[:, 1], x[:, 2]
    on = (np.abs(x[:, 0]) < 1e-12) & (y >= lo[0]) & (y <= hi[0]) & (z >= lo[1]) & (z <= hi[1])
    prof = np.zeros((len(x), 3))
    prof[on, 0] = (np.sin(np.pi * (y[on] - lo[0]) / (hi[0] - lo[0])) *
                   np.sin(np.pi * (z[on] - lo[1]) / (hi[1] - lo[1])))
    return prof, on


def quench_inflow(dims=(16, 16, 16), seed=5, amp=0.05, diffuse=0.0, rate=None, elem="hex"):
    """The quench tank with an inlet: cold liquid (phi = 0, 300 K) enters
    through a patch of the wall x = 0 with the normal velocity amp x
    ``inflow_bump`` and leaves through the open top; the inlet nodes hold
    their VOF value and temperature (IBC columns 8 and 6).  A per-node
    conductivity scale exp(0.1 z), z seeded standard normal, and the solid
    block's lumped nodal volumes (for the heat it gives off) ride along.

    ``diffuse`` > 0 replaces the sharp VOF field by the diffuse interface
    0.5 + 0.45 tanh((z - 0.6) / diffuse) (no bubble; every value inside (0, 1)
    away from the clamps); ``rate`` sets both phase-change coefficients
    (c_evap, c_cond; 0.1 in the tank's material); ``elem`` "tet" splits each
    hex into 6 linear tets, "prism" into 2 linear prisms (the node fields and
    conditions are unchanged)."""
    prob = quench(dims=dims, seed=seed)
    if elem == "tet":
        prob.mesh = M.hex_to_tets(prob.mesh)
    elif elem == "prism":
        prob.mesh = M.hex_to_prisms(prob.mesh)
    m = prob.mesh
    N = m.nnode
    if diffuse > 0.0:
        phi = 0.5 + 0.45 * np.tanh((m.xg[:, 2] - 0.6) / diffuse)
        for k in ("phig", "phigold"):
            prob.fields[k][:] = phi
    if rate is not None:
        prob.material.update(c_evap=float(rate), c_cond=float(rate))
    prof, on = inflow_bump(m.xg)
    prob.ibc[on, 5] = 1
    prob.ibc[on, 7] = 1
    prob.inflow_profile = prof
    prob.inflow_nodes = on
    prob.amp = float(amp)
    prob.bcval = amp * prof
    rng = np.random.default_rng(seed)
    prob.kscale = np.exp(0.1 * rng.standard_normal(N))
    prob.solid_volume = lumped_volumes(m, m.elm_id == 101)
    return prob


def _rotz(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def overset(nb=(96, 48, 48), nc=32, seed=1, rhoinf=0.5, theta=None) -> Problem:
    lo, hi = (0.0, 0.0, 0.0), (2.0, 1.0, 1.0)
    bg = M.structured_hex(*nb, lo=lo, hi=hi, elm_id=1)
    half = 0.2
    comp = M.structured_hex(nc, nc, nc, lo=(-half,) * 3, hi=(half,) * 3, elm_id=2)
    rng = np.random.default_rng(seed)
    if theta is None:
        theta = np.deg2rad(rng.uniform(0.0, 30.0))
    R = _rotz(theta)
    centre = np.array([0.9, 0.5, 0.5])
    xc = comp.xg @ R.T + centre
    Nb = bg.nnode
    mesh = M.Mesh(xg=np.concatenate([bg.xg, xc]),
                  ien=np.concatenate([bg.ien, comp.ien + Nb]).astype(np.int32),
                  elm_id=np.concatenate([bg.elm_id, comp.elm_id]).astype(np.int32))
    N = mesh.nnode
    # component outer boundary (local frame)
    comp_bnd = np.nonzero(np.any(np.abs(np.abs(comp.xg) - half) < 1e-12, axis=1))[0] + Nb
    # background fringe: inside the component, between 2 h_c and 2 h_c + h_b deep
    hc = 2 * half / nc
    hb = (hi[0] - lo[0]) / nb[0]
    loc = (bg.xg - centre) @ R          # background nodes in the component frame
    depth = half - np.abs(loc).max(axis=1)
    bg_fr = np.nonzero((depth >= 2 * hc) & (depth < 2 * hc + hb))[0]
    faces = M.boundary_masks(bg, lo, hi)
    ibc = np.zeros((N, 4), np.int32)
    inflow = np.zeros(N, bool)
    inflow[:Nb] = faces["xlo"]
    ibc[inflow, :3] = 1
    yw = np.zeros(N, bool)
    yw[:Nb] = faces["ylo"] | faces["yhi"]
    zw = np.zeros(N, bool)
    zw[:Nb] = faces["zlo"] | faces["zhi"]
    ibc[yw, 1] = 1
    ibc[zw, 2] = 1
    # the fringe encloses the component with velocity data only
    # (setBCs.f90:158-165), which leaves its pressure level free: it is held
    # at the node nearest the component's centre, as the cavity's is
    pin = Nb + int(np.argmin(np.linalg.norm(comp.xg, axis=1)))
    ibc[pin, 3] = 1
    bcval = np.zeros((N, 3))
    bcval[inflow, 0] = 1.0
    f = _zero_fields(N)
    u0 = np.zeros((N, 3))
    u0[:, 0] = 1.0
    u0[yw, 1] = 0.0
    u0[zw, 2] = 0.0
    f["ug"][:] = u0
    f["ugold"][:] = u0
    p = dict(genalpha(rhoinf), delt=0.005, rho=1.0, mu=8e-4, kdc=0.0,
             grav=(0.0, 0.0, -9.81))
    prob = Problem("overset", mesh, ibc, bcval, f, p)
    prob.fringe = {"component": (comp_bnd, 1), "background": (bg_fr, 2)}
    prob.pressure_pin = pin
    prob.theta = float(theta)
    return prob
