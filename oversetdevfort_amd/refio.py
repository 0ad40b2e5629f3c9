"""The reference's mesh and restart files, read and written on the host.

* ``part.<id>.dat`` (input_mesh, input.f90:15-135): NSD NSHL NSHLBmax; NNODE
  NELEM NBOUND NPATCH; the nodes (x, y, z[, NodeID]); the elements (ELMNSHL,
  IEN[, ELM_ID]); per boundary FACE_ID NFACE NNODE, the face node lists
  ([NSHLB,] FACE_IEN), the F2E FACE_OR pairs and BNODES; with NPATCH > 0 the
  NURBS weights and EPID/EIJK.  The bracketed items are present when
  fem_flag = 1 (NodeID, ELM_ID) or absent (NSHLB) -- the reference's choice.
* ``l2b.<id>.dat`` (input_mesh, input.f90:166-210; written by partmesh,
  genmesh_overset/partmesh.f90:476-510): per boundary FACE_ID NNODE NFACE,
  then per boundary node its volume node and its node in the unpartitioned
  boundary (shell) mesh (L2SNODE), and per face F2E and its shell face
  (L2SELEM); input_mesh checks FACE_ID, the counts and F2E.
* ``bmesh.<id>.dat`` (input_bmesh, input.f90:217-292; written by
  writeBndMesh, genmesh_overset/mesh.f90:211-252): NSD NSHLB; NNODE NFACE
  boundary-index FACE_ID; the node coordinates; the face node lists in the
  boundary's own numbering.
* ``restart.<step>.<rank>`` (writeSol, iosol.f90:377-454; readSol,
  iosol.f90:122-194): step and time, then dg, ug, ugm, acg, acgm (three per
  line), phig, rphig, pg, Tg, rTg (one per line), node by node.  writeSol
  writes the *old* temperature and its rate (Tgold, rTgold, iosol.f90:424-429)
  beside the current values of the other fields; readSol reads all ten into
  the old arrays.  ``write_restart`` takes the ten arrays as given, so a
  caller that wants the reference's file passes Tgold and rTgold.

Both are Fortran list-directed text.  The reader follows list-directed input:
each READ starts a new record and takes as many values as it needs, from the
following records if the record runs short, and drops the rest of its last
record; ``r*c`` repeats and D exponents are understood.  The writer prints
every real with 17 significant digits, so a file read back gives the same
bits.  Indices are 0-based on the Python side and 1-based in the files.
File names follow cname (cname.f90): "." and the decimal digits.
"""
from __future__ import annotations

import os

import numpy as np

from .mesh import Mesh

RESTART_FIELDS = ("dg", "ug", "ugm", "acg", "acgm", "phig", "rphig", "pg", "Tg", "rTg")
_VECTOR = {"dg", "ug", "ugm", "acg", "acgm"}


def cname(i: int) -> str:
    """cname.f90: "." followed by the decimal digits of i."""
    return "." + str(int(i))


def part_name(id_: int) -> str:
    return "part" + cname(id_) + ".dat"


def restart_name(step: int, rank: int = 1) -> str:
    return "restart" + cname(step) + cname(rank)


def l2b_name(id_: int) -> str:
    return "l2b" + cname(id_) + ".dat"


def bmesh_name(id_: int) -> str:
    return "bmesh" + cname(id_) + ".dat"


class _Records:
    """Fortran list-directed READ over the records (lines) of a text file."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def _tokens(self, line):
        out = []
        for t in line.replace(",", " ").split():
            if "*" in t:
                r, v = t.split("*", 1)
                out.extend([v] * int(r))
            else:
                out.append(t)
        return out

    def read(self, n):
        """the next n values of one READ statement, as strings"""
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise ValueError("end of file inside a READ")
            vals.extend(self._tokens(self.lines[self.pos]))
            self.pos += 1
        return vals[:n]

    def ints(self, n):
        return [int(v) for v in self.read(n)]

    def reals(self, n):
        return [float(v.replace("D", "E").replace("d", "e")) for v in self.read(n)]


def _real(v: float) -> str:
    return f"{v:.17g}" if np.isfinite(v) else repr(float(v))


def read_part(path: str, fem_flag: int = 1) -> dict:
    """input_mesh: returns mesh (Mesh, 0-based IEN), node_id, nshlbmax,
    bounds (list of dicts: face_id, face_ien (NFACE, NSHLBmax) 0-based, nshlb,
    f2e 0-based, face_or, bnodes 0-based), and wg/epid/eijk when NPATCH > 0."""
    with open(path) as f:
        r = _Records(f.read())
    nsd, nshl, nshlbmax = r.ints(3)
    nnode, nelem, nbound, npatch = r.ints(4)
    xg = np.zeros((nnode, nsd))
    node_id = np.zeros(nnode, np.int32)
    for i in range(nnode):
        if fem_flag == 1:
            v = r.read(nsd + 1)
            xg[i] = [float(t.replace("D", "E").replace("d", "e")) for t in v[:nsd]]
            node_id[i] = int(v[nsd])
        else:
            xg[i] = r.reals(nsd)
    ien = np.zeros((nelem, nshl), np.int32)
    elmnshl = np.zeros(nelem, np.int32)
    elm_id = np.zeros(nelem, np.int32)
    for e in range(nelem):
        # ELMNSHL(i) first, then as many nodes as it says, then ELM_ID
        first = r.read(1)
        r.pos -= 1
        k = int(first[0])
        v = r.ints(1 + k + (1 if fem_flag == 1 else 0))
        elmnshl[e] = v[0]
        ien[e, :k] = np.array(v[1:1 + k]) - 1
        if fem_flag == 1:
            elm_id[e] = v[1 + k]
    bounds = []
    for _ in range(nbound):
        face_id, nface, bnn = r.ints(3)
        face_ien = np.zeros((nface, nshlbmax), np.int32)
        nshlb = np.full(nface, nshlbmax, np.int32)
        for j in range(nface):
            if fem_flag == 1:
                face_ien[j] = np.array(r.ints(nshlbmax)) - 1
            else:
                v = r.ints(1 + nshlbmax)
                nshlb[j] = v[0]
                face_ien[j] = np.array(v[1:]) - 1
        f2e = np.zeros(nface, np.int32)
        face_or = np.zeros(nface, np.int32)
        for j in range(nface):
            a, b = r.ints(2)
            f2e[j], face_or[j] = a - 1, b
        bnodes = np.array([r.ints(1)[0] - 1 for _ in range(bnn)], np.int32)
        bounds.append(dict(face_id=face_id, face_ien=face_ien, nshlb=nshlb, f2e=f2e,
                           face_or=face_or, bnodes=bnodes))
    out = dict(mesh=Mesh(xg=xg, ien=ien, elm_id=elm_id), node_id=node_id, elmnshl=elmnshl,
               nsd=nsd, nshlbmax=nshlbmax, npatch=npatch, bounds=bounds)
    if npatch > 0:
        out["wg"] = np.array([r.reals(1)[0] for _ in range(nnode)])
        epid = np.zeros(nelem, np.int32)
        eijk = np.zeros((nelem, nsd), np.int32)
        for e in range(nelem):
            v = r.ints(1 + nsd)
            epid[e], eijk[e] = v[0], v[1:]
        out["epid"], out["eijk"] = epid, eijk
    return out


def write_part(path: str, mesh: Mesh, node_id=None, bounds=(), nshlbmax=None,
               fem_flag: int = 1) -> None:
    """The file input_mesh reads (no NURBS patches).  bounds: dicts with
    face_id, face_ien (NFACE, NSHLBmax) 0-based, f2e 0-based, face_or,
    bnodes 0-based (nshlb optional, for fem_flag 0)."""
    nnode, nelem, nshl = mesh.nnode, mesh.nelem, mesh.nshl
    if nshlbmax is None:
        nshlbmax = 4 if nshl == 8 else 3
    if node_id is None:
        node_id = np.zeros(nnode, np.int32)
    lines = [f"{3} {nshl} {nshlbmax}", f"{nnode} {nelem} {len(bounds)} 0"]
    for i in range(nnode):
        x = " ".join(_real(v) for v in mesh.xg[i])
        lines.append(f"{x} {int(node_id[i])}" if fem_flag == 1 else x)
    for e in range(nelem):
        nodes = " ".join(str(int(n) + 1) for n in mesh.ien[e])
        tail = f" {int(mesh.elm_id[e])}" if fem_flag == 1 else ""
        lines.append(f"{nshl} {nodes}{tail}")
    for b in bounds:
        fi = np.asarray(b["face_ien"])
        lines.append(f"{int(b['face_id'])} {fi.shape[0]} {len(b['bnodes'])}")
        for j in range(fi.shape[0]):
            nodes = " ".join(str(int(n) + 1) for n in fi[j])
            if fem_flag == 1:
                lines.append(nodes)
            else:
                nb = int(b.get("nshlb", np.full(fi.shape[0], nshlbmax))[j])
                lines.append(f"{nb} {nodes}")
        for e, o in zip(b["f2e"], b["face_or"]):
            lines.append(f"{int(e) + 1} {int(o)}")
        lines.extend(str(int(n) + 1) for n in b["bnodes"])
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_restart(path: str, nnode: int) -> dict:
    """readSol: step, time and the ten node fields (N,3) or (N,), in the
    file's order (dg, ug, ugm, acg, acgm, phig, rphig, pg, Tg, rTg)."""
    with open(path) as f:
        r = _Records(f.read())
    v = r.read(2)
    out = {"step": int(v[0]), "time": float(v[1].replace("D", "E").replace("d", "e"))}
    for name in RESTART_FIELDS:
        if name in _VECTOR:
            out[name] = np.array([r.reals(3) for _ in range(nnode)]).reshape(nnode, 3)
        else:
            out[name] = np.array([r.reals(1)[0] for _ in range(nnode)])
    return out


def write_restart(path: str, step: int, time: float, fields: dict) -> None:
    """writeSol's file from the ten arrays of RESTART_FIELDS (Tg and rTg as
    the caller gives them; the reference writes the old ones)."""
    lines = [f"{int(step)} {_real(time)}"]
    for name in RESTART_FIELDS:
        a = np.asarray(fields[name], np.float64)
        if name in _VECTOR:
            lines.extend(" ".join(_real(v) for v in row) for row in a.reshape(-1, 3))
        else:
            lines.extend(_real(v) for v in a.ravel())
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_l2b(path: str, bounds) -> None:
    """The l2b file of one partition: for each boundary its header, the
    (volume node, shell node) pair of each boundary node and the (element,
    shell face) pair of each face; l2snode and l2selem default to the
    boundary's own numbering (one partition)."""
    lines = []
    for b in bounds:
        bn = np.asarray(b["bnodes"])
        nface = len(b["f2e"])
        l2s = np.asarray(b.get("l2snode", np.arange(len(bn))))
        l2e = np.asarray(b.get("l2selem", np.arange(nface)))
        lines.append(f"{int(b['face_id'])} {len(bn)} {nface}")
        lines.extend(f"{int(n) + 1} {int(k) + 1}" for n, k in zip(bn, l2s))
        lines.extend(f"{int(e) + 1} {int(k) + 1}" for e, k in zip(b["f2e"], l2e))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_l2b(path: str, nbound: int) -> list:
    """input_mesh's reading of the l2b file: per boundary face_id, the volume
    nodes and L2SNODE, the elements and L2SELEM (all 0-based)."""
    with open(path) as f:
        r = _Records(f.read())
    out = []
    for _ in range(nbound):
        face_id, nn, nf = r.ints(3)
        vn, l2s = zip(*[r.ints(2) for _ in range(nn)]) if nn else ((), ())
        ve, l2e = zip(*[r.ints(2) for _ in range(nf)]) if nf else ((), ())
        out.append(dict(face_id=face_id, vnodes=np.array(vn, np.int32) - 1,
                        l2snode=np.array(l2s, np.int32) - 1, f2e=np.array(ve, np.int32) - 1,
                        l2selem=np.array(l2e, np.int32) - 1))
    return out


def write_bmesh(path: str, mesh: Mesh, bound: dict, index: int, nshlb: int = 4) -> None:
    """writeBndMesh's file for one boundary (index: its 1-based position)."""
    bn = np.asarray(bound["bnodes"])
    local = {int(n): i for i, n in enumerate(bn)}
    fi = np.asarray(bound["face_ien"])
    lines = [f"3 {nshlb}", f"{len(bn)} {fi.shape[0]} {int(index)} {int(bound['face_id'])}"]
    lines.extend(" ".join(_real(v) for v in mesh.xg[n]) for n in bn)
    lines.extend("".join(f"{local[int(n)] + 1:8d}" for n in row) for row in fi)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_bmesh(path: str) -> dict:
    """The boundary mesh: nsd, nshlb, nnode, nface, index, face_id, the node
    coordinates and the face node lists (0-based, the boundary's numbering)."""
    with open(path) as f:
        r = _Records(f.read())
    nsd, nshlb = r.ints(2)
    nnode, nface, index, face_id = r.ints(4)
    xg = np.array([r.reals(nsd) for _ in range(nnode)]).reshape(nnode, nsd)
    fi = np.array([r.ints(nshlb) for _ in range(nface)], np.int32).reshape(nface, nshlb) - 1
    return dict(nsd=nsd, nshlb=nshlb, nnode=nnode, nface=nface, index=index,
                face_id=face_id, xg=xg, face_ien=fi)


def bound_from_faces(mesh: Mesh, face_id: int, f2e, face_or) -> dict:
    """A boundary record for write_part from (f2e, face_or): the face node
    lists (hex faces in HEX_FACE_NODES order) and the sorted boundary nodes."""
    from .mesh import HEX_FACE_NODES
    f2e = np.asarray(f2e, np.int32)
    face_or = np.asarray(face_or, np.int32)
    fi = np.array([[mesh.ien[e, a] for a in HEX_FACE_NODES[o - 1]]
                   for e, o in zip(f2e, face_or)], np.int32).reshape(len(f2e), 4)
    return dict(face_id=face_id, face_ien=fi, f2e=f2e, face_or=face_or,
                bnodes=np.unique(fi).astype(np.int32))


def exists_part(dirname: str, id_: int) -> bool:
    return os.path.exists(os.path.join(dirname, part_name(id_)))
