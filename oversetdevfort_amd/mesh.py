"""Structured hex (and tet) meshes in the reference's in-memory conventions.

Host-side setup only: the reference reads meshes from ``part.<id>.dat``
(input.f90:15-135); this module builds the same arrays in memory.

* ``xg``   (NNODE, 3) float64, node coordinates.
* ``ien``  (NELEM, NSHL) int32, 0-based node ids.  Hex node order is the one of
  ``linhexshl`` (eval_shape_FEM.f90:161-168) in (xi, eta, zeta) = (x, y, z):
  1:(-,+,-) 2:(+,+,-) 3:(+,+,+) 4:(-,+,+) 5:(-,-,-) 6:(+,-,-) 7:(+,-,+) 8:(-,-,+).
* Node ids are lexicographic with x fastest: n = i + (nx+1)*(j + (ny+1)*k);
  element ids likewise over cells.
"""
from __future__ import annotations

import dataclasses

import numpy as np

# (di, dj, dk) offsets of the 8 hex nodes in linhexshl order
HEX_CORNERS = np.array(
    [[0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1],
     [0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], dtype=np.int64)


@dataclasses.dataclass
class Mesh:
    xg: np.ndarray          # (N, 3) float64
    ien: np.ndarray         # (E, NSHL) int32, 0-based
    elm_id: np.ndarray      # (E,) int32
    dims: tuple = ()        # (nx, ny, nz) cells for structured meshes

    @property
    def nnode(self) -> int:
        return int(self.xg.shape[0])

    @property
    def nelem(self) -> int:
        return int(self.ien.shape[0])

    @property
    def nshl(self) -> int:
        return int(self.ien.shape[1])


def node_index(i, j, k, nx, ny):
    return i + (nx + 1) * (j + (ny + 1) * k)


def structured_hex(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                   jitter=0.0, seed=1, elm_id=1) -> Mesh:
    """Box [lo, hi] split into nx*ny*nz trilinear hexes.

    ``jitter`` moves every interior node by U(-jitter*h, +jitter*h) per
    component (h = cell size in that direction), drawn from ``seed``.
    """
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    h = (hi - lo) / np.array([nx, ny, nz], dtype=np.float64)
    i = np.arange(nx + 1)
    j = np.arange(ny + 1)
    k = np.arange(nz + 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    xg = np.stack([lo[0] + I.ravel() * h[0],
                   lo[1] + J.ravel() * h[1],
                   lo[2] + K.ravel() * h[2]], axis=1).astype(np.float64)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=xg.shape) * h[None, :]
        interior = ((I.ravel() > 0) & (I.ravel() < nx) & (J.ravel() > 0) &
                    (J.ravel() < ny) & (K.ravel() > 0) & (K.ravel() < nz))
        xg[interior] += d[interior]
    ci = np.arange(nx)
    cj = np.arange(ny)
    ck = np.arange(nz)
    CK, CJ, CI = np.meshgrid(ck, cj, ci, indexing="ij")
    CI, CJ, CK = CI.ravel(), CJ.ravel(), CK.ravel()
    ien = np.empty((CI.size, 8), dtype=np.int32)
    for a, (di, dj, dk) in enumerate(HEX_CORNERS):
        ien[:, a] = node_index(CI + di, CJ + dj, CK + dk, nx, ny)
    return Mesh(xg=xg, ien=ien, elm_id=np.full(ien.shape[0], elm_id, np.int32),
                dims=(nx, ny, nz))


def hex_colors(mesh: Mesh) -> np.ndarray:
    """8-colouring of a structured hex mesh: elements of one colour share no node."""
    nx, ny, nz = mesh.dims
    e = np.arange(mesh.nelem)
    i = e % nx
    j = (e // nx) % ny
    k = e // (nx * ny)
    return ((i & 1) + 2 * (j & 1) + 4 * (k & 1)).astype(np.int32)


def hex_to_tets(mesh: Mesh) -> Mesh:
    """Split each hex into 6 Kuhn tets sharing the cell diagonal (0,0,0)-(1,1,1).

    Tet node order follows ``lintetshl`` (N1=xi, N2=eta, N3=zeta,
    N4=1-xi-eta-zeta; eval_shape_FEM.f90:201-225): node 4 is the vertex the
    local frame starts from.  Each tet is made positively oriented.  The
    hexes need not come from one structured block (the overset pair's two
    blocks split the same way).
    """
    # local corner id -> linhexshl slot
    slot = {tuple(c): a for a, c in enumerate(HEX_CORNERS.tolist())}
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        c0 = [0, 0, 0]
        c1 = list(c0)
        c1[p[0]] = 1
        c2 = list(c1)
        c2[p[1]] = 1
        c3 = [1, 1, 1]
        tets.append([slot[tuple(c1)], slot[tuple(c2)], slot[tuple(c3)],
                     slot[tuple(c0)]])
    out = []
    for t in tets:
        out.append(mesh.ien[:, t])
    ien = np.stack(out, axis=1).reshape(-1, 4).astype(np.int32)
    # orientation: det[x1-x4, x2-x4, x3-x4] > 0
    x = mesh.xg
    d = np.linalg.det(np.stack([x[ien[:, 0]] - x[ien[:, 3]],
                                x[ien[:, 1]] - x[ien[:, 3]],
                                x[ien[:, 2]] - x[ien[:, 3]]], axis=2))
    neg = d < 0
    ien[neg, 0], ien[neg, 1] = ien[neg, 1].copy(), ien[neg, 0].copy()
    elm = np.repeat(mesh.elm_id, 6)
    return Mesh(xg=mesh.xg.copy(), ien=ien, elm_id=elm.astype(np.int32),
                dims=mesh.dims)


def hex_to_prisms(mesh: Mesh) -> Mesh:
    """Split each hex into 2 linear prisms (wedges) along the diagonal of its
    z faces.  Prism node order follows ``linprishl`` (eval_shape_FEM.f90:
    438-477): nodes 1-3 the bottom triangle (zeta = -1) with node 1 at xi = 1,
    node 2 at eta = 1, node 3 at the origin of (xi, eta); nodes 4-6 the same
    corners on the top.  Each prism is made positively oriented."""
    slot = {tuple(c): a for a, c in enumerate(HEX_CORNERS.tolist())}
    tris = [((1, 0), (0, 1), (0, 0)), ((0, 1), (1, 0), (1, 1))]
    out = []
    for t in tris:
        idx = [slot[(i, j, 0)] for i, j in t] + [slot[(i, j, 1)] for i, j in t]
        out.append(mesh.ien[:, idx])
    ien = np.stack(out, axis=1).reshape(-1, 6).astype(np.int32)
    # orientation: dx/dxi of linprishl at the centroid; flip nodes 1 <-> 2 and
    # 4 <-> 5 where its determinant is negative
    x = mesh.xg
    c = x[ien]
    e1 = 0.5 * ((c[:, 0] - c[:, 2]) + (c[:, 3] - c[:, 5]))
    e2 = 0.5 * ((c[:, 1] - c[:, 2]) + (c[:, 4] - c[:, 5]))
    e3 = (c[:, 3:].mean(axis=1) - c[:, :3].mean(axis=1))
    d = np.linalg.det(np.stack([e1, e2, e3], axis=2))
    neg = d < 0
    ien[neg, 0], ien[neg, 1] = ien[neg, 1].copy(), ien[neg, 0].copy()
    ien[neg, 3], ien[neg, 4] = ien[neg, 4].copy(), ien[neg, 3].copy()
    elm = np.repeat(mesh.elm_id, 2)
    return Mesh(xg=mesh.xg.copy(), ien=ien, elm_id=elm.astype(np.int32), dims=mesh.dims)


# Local nodes of each hex face, FACE_OR 1..6 = xi -1, xi +1, eta -1, eta +1,
# zeta -1, zeta +1 (the reference has no hex face map; this numbering is the
# one of the face kernels and the oracle).
HEX_FACE_NODES = tuple(
    tuple(int(a) for a in np.nonzero(HEX_CORNERS[:, c] == side)[0])
    for c in range(3) for side in (0, 1))

# FACE_OR of a tet face by the local node it leaves out (genGPtritet and
# eval_faceshape_tet, genGP_GW.f90:478-572, eval_shape_FEM.f90:84-99: face 1
# holds nodes 2,3,1; 2: 4,3,2; 3: 1,3,4; 4: 2,1,4)
TET_FACE_BY_MISSING = {3: 1, 0: 2, 1: 3, 2: 4}


def boundary_faces(mesh: Mesh, node_mask=None):
    """(f2e, face_or): 0-based element and FACE_OR of every face that no
    other element shares, sorted by (element, face); with ``node_mask`` only
    the faces whose nodes all lie in it."""
    if mesh.nshl == 8:
        local = [(fo + 1, HEX_FACE_NODES[fo]) for fo in range(6)]
    elif mesh.nshl == 4:
        local = [(TET_FACE_BY_MISSING[miss], tuple(a for a in range(4) if a != miss))
                 for miss in range(4)]
    else:
        raise ValueError("boundary faces of hex and tet meshes only")
    seen = {}
    for e, nodes in enumerate(np.asarray(mesh.ien)):
        for fo, loc in local:
            key = tuple(sorted(int(nodes[a]) for a in loc))
            seen.setdefault(key, []).append((e, fo))
    faces = []
    for key, v in seen.items():
        if len(v) == 1 and (node_mask is None or all(node_mask[n] for n in key)):
            faces.append(v[0])
    faces.sort()
    return (np.array([f[0] for f in faces], np.int32),
            np.array([f[1] for f in faces], np.int32))


def box_faces(mesh: Mesh, sides):
    """(f2e, face_or) of the faces of a structured box mesh on the given
    sides ('xlo', 'xhi', 'ylo', 'yhi', 'zlo', 'zhi'), sorted by (element,
    face); the same faces as boundary_faces with the side's node mask."""
    nx, ny, nz = mesh.dims
    e = np.arange(nx * ny * nz)
    ci, cj, ck = e % nx, (e // nx) % ny, e // (nx * ny)
    pick = {"xlo": (ci == 0, 1), "xhi": (ci == nx - 1, 2), "ylo": (cj == 0, 3),
            "yhi": (cj == ny - 1, 4), "zlo": (ck == 0, 5), "zhi": (ck == nz - 1, 6)}
    f2e = np.concatenate([e[pick[s][0]] for s in sides]).astype(np.int32)
    fo = np.concatenate([np.full(int(pick[s][0].sum()), pick[s][1]) for s in sides])
    order = np.lexsort((fo, f2e))
    return f2e[order], fo[order].astype(np.int32)


def boundary_masks(mesh: Mesh, lo, hi, tol=1e-12):
    """Boolean node masks of the six faces of a box mesh: xlo,xhi,ylo,yhi,zlo,zhi."""
    x = mesh.xg
    out = {}
    for d, name in enumerate("xyz"):
        out[name + "lo"] = np.abs(x[:, d] - lo[d]) < tol
        out[name + "hi"] = np.abs(x[:, d] - hi[d]) < tol
    return out
