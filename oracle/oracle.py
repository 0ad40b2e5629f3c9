"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker (or, for the
baseline, as the thing timed on the host cores).  The HIP product path never
imports it and fails loudly when its own library is missing.

Pinning of each oracle function (what checks it independently of itself):

=====================  =====================================================
function               pinned by
=====================  =====================================================
genalpha               closed form (param.f90:469-480) + reference harness
gauss_volume           reference harness genGPandGWtetpri bit-exact;
                       exact integration of monomials (tests)
linhexshl/tet/pri      reference harness bit-exact; partition of unity,
                       Kronecker property, gradient sums (tests)
inverse3x3             reference harness bit-exact; A @ Ainv = I
eval_shape_fem         reference harness (tet, prism); affine volume
                       identity sum DetJ*w = |element| for hex (tests)
sparsity               reference genSparStruc bit-exact; brute force set
                       construction; closed form prod(3n-2) (tests)
assemble_ns            reference IntElmAss_NS_overset on tet meshes
                       (element kernels, BC, scatter: <=1e-13); hex differs
                       only by the shape/quadrature pinned above
spmv_up                reference SparseProdUP_3D; dense product (tests)
gmres_up               reference SparseGMRES_up; Arnoldi invariants and a
                       direct solve (tests)
interp_tet             reference interpolation_tet values + donors parsed
                       from its own "detected element" print; brute force
interp_hex             (extension, no reference code) closed-form local
                       coordinates on axis-aligned and affine (rotated)
                       grids; brute-force first-index donor over all
                       elements; exact interpolation of linear fields
assemble_tem           reference IntElmAss_Tem_Quenching_STAB + BCLhs_tem +
                       FillSparseMat_tem on tet meshes of fluid and solid
                       elements: bit-exact, and within one rounding of
                       norm2 with discontinuity capturing; Dirichlet rows
                       are identity rows with zero RHS (tests)
spmv_s                 reference SparseProd_LS bit-exact
gmres_tem              reference SparseGMRES_tem bit-exact (solution)
assemble_nsvof         reference IntElmAss_NSVOF_Quenching_STAB relinked to
                       BCLhs_NS_conv + FillSparseMat_NS_conv on tet meshes:
                       all eight LHS arrays and three RHS bit-exact, and
                       within one rounding of norm2 with VOF discontinuity
                       capturing
spmv_ns_conv           reference SparseProd_NS_conv bit-exact
gmres_ns_conv          reference SparseGMRES_ns_conv bit-exact (solution)
assemble_ns_col        the same per-element routine as assemble_ns in colour
                       order; equal to assemble_ns to rounding (tests)
=====================  =====================================================

``using(path)`` swaps in another build of the same source for the host
baseline (``make -C oracle native``: -O3 -march=native with OpenMP), whose
rounding differs from the parity build's.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        _LIB = _load(path)
    return _LIB


class using:
    """``with using(path):`` every wrapper calls the build at ``path``."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _LIB
        self.saved = _LIB
        _LIB = _load(self.path)
        return _LIB

    def __exit__(self, *a):
        global _LIB
        _LIB = self.saved


def _load(path):
    """The oracle library at ``path`` with every signature set."""
    L = ctypes.CDLL(path)
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_threads.restype = ctypes.c_int
    L.orc_assemble_ns_col.argtypes = (
        [ctypes.c_int] * 4 + [_ip, _ip, _ip, ctypes.c_int, ctypes.c_int, _dp, _ip] +
        [_dp] * 7 + [_ip, _ip, ctypes.c_int, _ip, _ip] + [_dp] * 7)
    L.orc_assemble_ns_col.restype = ctypes.c_int
    L.orc_genalpha.argtypes = [ctypes.c_double, _dp]
    L.orc_gauss_volume.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_gauss_volume.restype = ctypes.c_int
    for f in (L.orc_linhexshl, L.orc_lintetshl, L.orc_linprishl):
        f.argtypes = [_dp, _dp, _dp]
    L.orc_inverse3x3.argtypes = [_dp, _dp, _dp]
    L.orc_eval_shape_fem.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp, _dp]
    L.orc_eval_shape_fem.restype = ctypes.c_int
    L.orc_stab_ns.argtypes = [_dp, ctypes.c_double, _dp, ctypes.c_double,
                              ctypes.c_double, _dp, _dp, _dp]
    L.orc_sparsity.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               _ip, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_sparsity.restype = ctypes.c_long
    L.orc_assemble_ns.argtypes = (
        [ctypes.c_int] * 4 + [_ip, _ip, _ip, ctypes.c_int, ctypes.c_int, _dp,
                              _ip, ctypes.c_int] + [_dp] * 6 + [ctypes.c_void_p, _dp,
                                                               ctypes.c_int] +
        [_ip, _ip] + [_dp] * 6 + [_dp])
    L.orc_assemble_ns.restype = ctypes.c_int
    L.orc_face_ns.argtypes = ([ctypes.c_int] * 3 + [_ip] * 3 + [ctypes.c_int] * 2 +
                              [_dp, _ip] + [_dp] * 5 + [ctypes.c_int, _ip, _ip,
                                                        ctypes.c_int, _ip, _ip] +
                              [_dp] * 6)
    L.orc_element_ns.argtypes = ([ctypes.c_int, ctypes.c_int] + [_dp] * 7 +
                                 [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int] +
                                 [_dp] * 7)
    L.orc_element_ns.restype = ctypes.c_int
    L.orc_bclhs_3d.argtypes = [ctypes.c_int, _ip] + [_dp] * 6
    L.orc_spmv_up.argtypes = [ctypes.c_int, _ip, _ip] + [_dp] * 8
    L.orc_gmres_up.argtypes = ([ctypes.c_int, _ip, _ip] + [_dp] * 8 +
                               [ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                _dp, _dp])
    L.orc_gmres_up.restype = ctypes.c_int
    L.orc_interp_tet.argtypes = ([ctypes.c_int] * 3 + [_ip, _ip, ctypes.c_int] +
                                 [_dp] * 3 + [ctypes.c_int, _dp, _dp, _ip, _dp])
    L.orc_assemble_tem.argtypes = ([ctypes.c_int] * 4 + [_ip, _ip, ctypes.c_int, _dp,
                                    _ip] + [_dp] * 6 + [ctypes.c_void_p] + [_dp] * 2 +
                                   [_ip, _ip, _dp, _dp])
    L.orc_assemble_tem.restype = ctypes.c_int
    L.orc_spmv_s.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
    L.orc_gmres_tem.argtypes = ([ctypes.c_int, _ip, _ip, _dp, _dp, _dp, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, _dp])
    L.orc_gmres_tem.restype = ctypes.c_int
    L.orc_assemble_nsvof.argtypes = ([ctypes.c_int] * 4 + [_ip, _ip, ctypes.c_int, _dp,
                                      _ip] + [_dp] * 11 + [_ip, _ip] + [_dp] * 12)
    L.orc_assemble_nsvof.restype = ctypes.c_int
    L.orc_spmv_ns_conv.argtypes = [ctypes.c_int, _ip, _ip] + [_dp] * 14
    L.orc_gmres_ns_conv.argtypes = ([ctypes.c_int, _ip, _ip] + [_dp] * 14 +
                                    [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp])
    L.orc_gmres_ns_conv.restype = ctypes.c_int
    L.orc_interp_hex.argtypes = ([ctypes.c_int] * 2 + [_ip, _ip, ctypes.c_int] +
                                 [_dp] * 3 + [ctypes.c_int, _dp, _dp, _ip, _dp])
    L.orc_hex_local.argtypes = [_dp, _dp, _dp]
    L.orc_hex_local.restype = ctypes.c_int
    return L


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _soa(a):
    """(N,3) node field -> column-major flat (3*N) as the reference stores it."""
    return _c(np.asarray(a, dtype=np.float64).T).ravel()


def genalpha(rhoinf):
    out = np.zeros(4)
    lib().orc_genalpha(float(rhoinf), out)
    return out


def gauss_volume(ngp):
    gp = np.zeros(3 * ngp)
    gw = np.zeros(ngp)
    if lib().orc_gauss_volume(ngp, gp, gw) != 0:
        raise ValueError(ngp)
    return gp.reshape(3, ngp).T.copy(), gw


def shape_ref(nshl, pt):
    f = {8: lib().orc_linhexshl, 4: lib().orc_lintetshl, 6: lib().orc_linprishl}[nshl]
    shl = np.zeros(nshl)
    sg = np.zeros(3 * nshl)
    f(_c(pt), shl, sg)
    return shl, sg.reshape(3, nshl).T.copy()


def inverse3x3(A):
    Ai = np.zeros(9)
    dj = np.zeros(1)
    lib().orc_inverse3x3(_c(np.asarray(A).T).ravel(), Ai, dj)
    return Ai.reshape(3, 3).T.copy(), float(dj[0])


def eval_shape_fem(nshl, pt, xl, dl):
    shlu = np.zeros(nshl)
    sg = np.zeros(3 * nshl)
    dxidx, G, Gi, dj = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(1)
    lib().orc_eval_shape_fem(nshl, _c(pt), _soa(xl), _soa(dl), shlu, sg, dxidx,
                             G, Gi, dj)
    return (shlu, sg.reshape(3, nshl).T.copy(), dxidx.reshape(3, 3).T.copy(),
            G.reshape(3, 3).T.copy(), Gi.reshape(3, 3).T.copy(), float(dj[0]))


def sparsity(nnode, ien):
    """genSparStruc -> (col[N+1], row[icnt]) 0-based."""
    ien = np.asarray(ien, dtype=np.int32)
    nelem, nshl = ien.shape
    ienf = _c(ien.T, np.int32).ravel()
    icnt = lib().orc_sparsity(nnode, nelem, nshl, ienf, None, None)
    if icnt < 0:
        raise RuntimeError("sparsity overflow")
    col = np.zeros(nnode + 1, np.int32)
    row = np.zeros(icnt, np.int32)
    lib().orc_sparsity(nnode, nelem, nshl, ienf, col.ctypes.data, row.ctypes.data)
    return col, row


def params_array(p):
    """p: dict with almi alfi gami beti delt rho mu kdc grav(3)."""
    g = p.get("grav", (0.0, 0.0, -9.81))
    return np.array([p["almi"], p["alfi"], p["gami"], p["beti"], p["delt"],
                     p["rho"], p["mu"], p.get("kdc", 0.0), g[0], g[1], g[2]],
                    dtype=np.float64)


def assemble_ns(mesh, ibc, state, prm, col, row, tensor_flag=6,
                subdomain_flag=1, elm_id_set=(1, 2), ngauss=None, mus=None, vms=0):
    """IntElmAss_NS_overset on zeroed outputs.

    state: dict of alpha-stage node fields dg, ug, ugm, acg, acgm (N,3), pg (N,).
    mus: optional per-node viscosity scale (N,) (an extension; None is the
    reference's constant viscosity).  vms=1: the residual with the fine-scale
    velocity uprime = -tauM rLi (an extension; 0 is the live call's).
    ibc: (N, 4) int.  Returns dict with k11 (icnt,9), g (icnt,3), d1 (icnt,3),
    m (icnt,), rhsu (N,3), rhsp (N,), cfl, nneg.
    """
    N, E, nshl = mesh.nnode, mesh.nelem, mesh.nshl
    if ngauss is None:
        ngauss = nshl
    icnt = row.size
    k11 = np.zeros(icnt * 9)
    g = np.zeros(icnt * 3)
    d1 = np.zeros(icnt * 3)
    m = np.zeros(icnt)
    rhsu = np.zeros(3 * N)
    rhsp = np.zeros(N)
    cfl = np.zeros(1)
    ibcf = _c(np.asarray(ibc)[:, :4].T, np.int32).ravel()
    mus_ = None if mus is None else _c(mus)    # kept alive across the call
    nneg = lib().orc_assemble_ns(
        N, E, nshl, ngauss, _c(mesh.ien.T, np.int32).ravel(),
        _c(mesh.elm_id, np.int32), _c(elm_id_set, np.int32), subdomain_flag,
        tensor_flag, _soa(mesh.xg), ibcf, 4, _soa(state["dg"]),
        _soa(state["ug"]), _soa(state["ugm"]), _soa(state["acg"]),
        _soa(state["acgm"]), _c(state["pg"]),
        None if mus_ is None else mus_.ctypes.data, _c(prm), int(vms), _c(col, np.int32),
        _c(row, np.int32), k11, g, d1, m, rhsu, rhsp, cfl)
    if nneg < 0:
        raise RuntimeError("oracle assembly failed")
    return dict(k11=k11.reshape(icnt, 9), g=g.reshape(icnt, 3),
                d1=d1.reshape(icnt, 3), m=m, rhsu=rhsu.reshape(3, N).T.copy(),
                rhsp=rhsp, cfl=float(cfl[0]), nneg=int(nneg))


def assemble_ns_col(mesh, ibc, state, prm, col, row, colors, tensor_flag=6,
                    subdomain_flag=1, elm_id_set=(1, 2)):
    """IntElmAss_NS_overset colour by colour (``colors``: colour of each
    element); the elements of a colour in parallel in an OpenMP build."""
    N, E, nshl = mesh.nnode, mesh.nelem, mesh.nshl
    colors = np.asarray(colors)
    order = np.argsort(colors, kind="stable").astype(np.int32)
    ncol = int(colors.max()) + 1
    off = np.zeros(ncol + 1, np.int32)
    off[1:] = np.cumsum(np.bincount(colors, minlength=ncol))
    icnt = row.size
    out = dict(k11=np.zeros(icnt * 9), g=np.zeros(icnt * 3), d1=np.zeros(icnt * 3),
               m=np.zeros(icnt), rhsu=np.zeros(3 * N), rhsp=np.zeros(N))
    cfl = np.zeros(1)
    ibcf = _c(np.asarray(ibc)[:, :4].T, np.int32).ravel()
    nneg = lib().orc_assemble_ns_col(
        N, E, nshl, nshl, _c(mesh.ien.T, np.int32).ravel(), _c(mesh.elm_id, np.int32),
        _c(elm_id_set, np.int32), subdomain_flag, tensor_flag, _soa(mesh.xg), ibcf,
        _soa(state["dg"]), _soa(state["ug"]), _soa(state["ugm"]), _soa(state["acg"]),
        _soa(state["acgm"]), _c(state["pg"]), _c(prm), _c(col, np.int32), _c(row, np.int32),
        ncol, off, order, out["k11"], out["g"], out["d1"], out["m"], out["rhsu"],
        out["rhsp"], cfl)
    if nneg < 0:
        raise RuntimeError("oracle assembly failed")
    return dict(k11=out["k11"].reshape(icnt, 9), g=out["g"].reshape(icnt, 3),
                d1=out["d1"].reshape(icnt, 3), m=out["m"],
                rhsu=out["rhsu"].reshape(3, N).T.copy(), rhsp=out["rhsp"], cfl=float(cfl[0]),
                nneg=int(nneg))


def face_ns(mesh, ibc, state, prm, col, row, faces, tensor_flag=6, subdomain_flag=1,
            elm_id_set=(1, 2), ngb=None, out=None):
    """FaceAssembly_NS_weak + FaceAssembly_NS_outflow over `faces` = (f2e
    0-based elements, face_or), added into `out` (a dict as assemble_ns
    returns; zeros if None).  ngb: points per face (hex 4, tet 1 or 3)."""
    N, E, nshl = mesh.nnode, mesh.nelem, mesh.nshl
    if ngb is None:
        ngb = 4 if nshl == 8 else 3
    icnt = row.size
    if out is None:
        out = dict(k11=np.zeros((icnt, 9)), g=np.zeros((icnt, 3)), d1=np.zeros((icnt, 3)),
                   m=np.zeros(icnt), rhsu=np.zeros((N, 3)), rhsp=np.zeros(N))
    k11 = _c(out["k11"]).ravel().copy()
    g = _c(out["g"]).ravel().copy()
    d1 = _c(out["d1"]).ravel().copy()
    m = _c(out["m"]).copy()
    rhsu = _c(np.asarray(out["rhsu"]).T).ravel().copy()
    rhsp = _c(out["rhsp"]).copy()
    f2e, fo = (np.ascontiguousarray(a, np.int32) for a in faces)
    ibcf = _c(np.asarray(ibc)[:, :4].T, np.int32).ravel()
    r = lib().orc_face_ns(
        N, E, nshl, _c(mesh.ien.T, np.int32).ravel(), _c(mesh.elm_id, np.int32),
        _c(elm_id_set, np.int32), subdomain_flag, tensor_flag, _soa(mesh.xg), ibcf,
        _soa(state["dg"]), _soa(state["ug"]), _soa(state["ugm"]), _c(state["pg"]), _c(prm),
        len(f2e), f2e, fo, int(ngb), _c(col, np.int32), _c(row, np.int32),
        k11, g, d1, m, rhsu, rhsp)
    if r < 0:
        raise RuntimeError("oracle face assembly failed")
    return dict(k11=k11.reshape(icnt, 9), g=g.reshape(icnt, 3), d1=d1.reshape(icnt, 3), m=m,
                rhsu=rhsu.reshape(3, N).T.copy(), rhsp=rhsp)


def spmv_up(col, row, A, xu, xp):
    N = col.size - 1
    yu = np.zeros(3 * N)
    yp = np.zeros(N)
    lib().orc_spmv_up(N, _c(col, np.int32), _c(row, np.int32),
                      _c(A["k11"]).ravel(), _c(A["g"]).ravel(), _c(A["d1"]).ravel(),
                      _c(A["m"]), _soa(xu), _c(xp), yu, yp)
    return yu.reshape(3, N).T.copy(), yp


def gmres_up(col, row, A, rhsu, rhsp, tol, kspace, kspace_min=0):
    N = col.size - 1
    solu = np.zeros(3 * N)
    solp = np.zeros(N)
    res = np.zeros(1)
    r0 = np.zeros(1)
    its = lib().orc_gmres_up(N, _c(col, np.int32), _c(row, np.int32),
                             _c(A["k11"]).ravel(), _c(A["g"]).ravel(),
                             _c(A["d1"]).ravel(), _c(A["m"]), _soa(rhsu),
                             _c(rhsp), solu, solp, float(tol), int(kspace),
                             int(kspace_min), res, r0)
    return dict(solu=solu.reshape(3, N).T.copy(), solp=solp, iters=int(its),
                resid=float(res[0]), rnorm0=float(r0[0]))


def _qvec(qprm):
    """The 20 material values of the C restatement (the 19 of the quench
    material, then the fluid id, 102 when absent); entries beyond are
    ignored (lsc_kdc goes in through ``ex``)."""
    qp = np.asarray(qprm, np.float64)[:20]
    q = np.zeros(20)
    q[:qp.size] = qp
    if qp.size < 20:
        q[19] = 102
    return q


def assemble_tem(mesh, ibc_t, st, prm, qprm, col, row, tensor_flag, lhs=None, rhs=None,
                 ks=None):
    """IntElmAss_Tem_Quenching_STAB + BCLhs_tem + FillSparseMat_tem.  st holds
    the alpha-stage fields dg, ug, ugm (N,3) and phig, Tg, rTg (N); prm the
    NS parameter vector (almi, alfi, gami, [beti], delt, ...) or the 4-vector
    (almi, alfi, gami, delt); qprm the 19 material values (+ fluid id); ks
    an optional per-node conductivity scale (an extension)."""
    N, E = mesh.nnode, mesh.nelem
    p = np.asarray(prm, np.float64)
    if p.size == 4:
        p = np.array([p[0], p[1], p[2], 0.0, p[3]])
    q = _qvec(qprm)
    lhs = np.zeros(len(row)) if lhs is None else lhs
    rhs = np.zeros(N) if rhs is None else rhs
    ks_ = None if ks is None else _c(ks)
    nneg = lib().orc_assemble_tem(
        N, E, mesh.nshl, mesh.nshl, _c(mesh.ien.T, np.int32).ravel(),
        _c(mesh.elm_id, np.int32), int(tensor_flag), _soa(mesh.xg),
        _c(ibc_t, np.int32), _soa(st["dg"]), _soa(st["ug"]), _soa(st["ugm"]),
        _c(st["phig"]), _c(st["Tg"]), _c(st["rTg"]),
        None if ks_ is None else ks_.ctypes.data_as(ctypes.c_void_p), _c(p), _c(q),
        _c(col, np.int32), _c(row, np.int32), lhs, rhs)
    return dict(lhs=lhs, rhs=rhs, nneg=nneg)


NSVOF_BLOCKS = ("k11", "g", "d1", "m", "ls", "lsu", "pls", "uls")
_NSVOF_WIDTH = dict(k11=9, g=3, d1=3, m=1, ls=1, lsu=3, pls=1, uls=3)


def assemble_nsvof(mesh, ibc8, st, prm, qprm, ex, col, row, tensor_flag):
    """IntElmAss_NSVOF_Quenching_STAB + BCLhs_NS_conv + FillSparseMat_NS_conv.
    st: alpha-stage dg, ug, ugm, acg (N,3) and pg, phig, rphig, Tg (N); prm as
    assemble_tem; ex = (gx, gy, gz, lsc_kdc); ibc8 (N, 8)."""
    N, E = mesh.nnode, mesh.nelem
    p = np.asarray(prm, np.float64)
    if p.size == 4:
        p = np.array([p[0], p[1], p[2], 0.0, p[3]])
    q = _qvec(qprm)
    Z = len(row)
    out = {k: np.zeros(Z * w) for k, w in _NSVOF_WIDTH.items()}
    rhsu = np.zeros(3 * N)
    rhsp = np.zeros(N)
    rhsl = np.zeros(N)
    cfl = np.zeros(1)
    nneg = lib().orc_assemble_nsvof(
        N, E, mesh.nshl, mesh.nshl, _c(mesh.ien.T, np.int32).ravel(),
        _c(mesh.elm_id, np.int32), int(tensor_flag), _soa(mesh.xg),
        _c(np.asarray(ibc8, np.int32).T, np.int32).ravel(), _soa(st["dg"]), _soa(st["ug"]),
        _soa(st["ugm"]), _soa(st["acg"]), _c(st["pg"]), _c(st["phig"]), _c(st["rphig"]),
        _c(st["Tg"]), _c(p), _c(q), _c(ex), _c(col, np.int32), _c(row, np.int32),
        *[out[k] for k in NSVOF_BLOCKS], rhsu, rhsp, rhsl, cfl)
    res = {k: (v.reshape(Z, _NSVOF_WIDTH[k]) if _NSVOF_WIDTH[k] > 1 else v)
           for k, v in out.items()}
    res.update(rhsu=rhsu.reshape(3, N).T.copy(), rhsp=rhsp, rhsls=rhsl, nneg=nneg,
               cfl=float(cfl[0]))
    return res


def _nsvof_mats(A):
    return [_c(A[k]).ravel() for k in NSVOF_BLOCKS]


def spmv_ns_conv(col, row, A, xu, xp, xl):
    N = len(col) - 1
    yu = np.zeros(3 * N)
    yp = np.zeros(N)
    yl = np.zeros(N)
    lib().orc_spmv_ns_conv(N, _c(col, np.int32), _c(row, np.int32), *_nsvof_mats(A),
                           _soa(xu), _c(xp), _c(xl), yu, yp, yl)
    return yu.reshape(3, N).T.copy(), yp, yl


def gmres_ns_conv(col, row, A, bu, bp, bl, tol, kspace, kspace_min):
    N = len(col) - 1
    su = np.zeros(3 * N)
    sp = np.zeros(N)
    sl = np.zeros(N)
    res = np.zeros(1)
    it = lib().orc_gmres_ns_conv(N, _c(col, np.int32), _c(row, np.int32), *_nsvof_mats(A),
                                 _soa(bu), _c(bp), _c(bl), su, sp, sl, float(tol),
                                 int(kspace), int(kspace_min), res)
    return dict(solu=su.reshape(3, N).T.copy(), solp=sp, solls=sl, iters=it,
                resid=float(res[0]))


def spmv_s(col, row, A, x):
    y = np.zeros(len(col) - 1)
    lib().orc_spmv_s(len(col) - 1, _c(col, np.int32), _c(row, np.int32), _c(A), _c(x), y)
    return y


def gmres_tem(col, row, A, b, tol, kspace, kspace_min):
    sol = np.zeros(len(col) - 1)
    res = np.zeros(1)
    it = lib().orc_gmres_tem(len(col) - 1, _c(col, np.int32), _c(row, np.int32), _c(A),
                             _c(b), sol, float(tol), int(kspace), int(kspace_min), res)
    return dict(sol=sol, iters=it, resid=float(res[0]))


def interp_hex(mesh, src_id, dg, ug, xr, dr):
    """Hex donor search + trilinear interpolation (extension of
    interpolation_tet to hexes)."""
    N, E = mesh.nnode, mesh.nelem
    nr = xr.shape[0]
    donor = np.zeros(nr, np.int32)
    val = np.zeros(3 * nr)
    lib().orc_interp_hex(N, E, _c(mesh.ien.T, np.int32).ravel(),
                         _c(mesh.elm_id, np.int32), int(src_id), _soa(mesh.xg),
                         _soa(dg), _soa(ug), nr, _soa(xr), _soa(dr), donor, val)
    return donor, val.reshape(3, nr).T.copy()


def hex_local(X, p):
    """(inside, xi) of point p in the hex with corners X (8,3)."""
    xi = np.zeros(3)
    ok = lib().orc_hex_local(_c(X, np.float64).ravel(), _c(p, np.float64), xi)
    return bool(ok), xi


def interp_tet(mesh, src_id, dg, ug, xr, dr):
    N, E = mesh.nnode, mesh.nelem
    nr = xr.shape[0]
    donor = np.zeros(nr, np.int32)
    val = np.zeros(3 * nr)
    lib().orc_interp_tet(N, E, mesh.nshl, _c(mesh.ien.T, np.int32).ravel(),
                         _c(mesh.elm_id, np.int32), int(src_id), _soa(mesh.xg),
                         _soa(dg), _soa(ug), nr, _soa(xr), _soa(dr), donor, val)
    return donor, val.reshape(3, nr).T.copy()
