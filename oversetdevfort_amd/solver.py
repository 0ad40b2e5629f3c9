"""Host-side mirror of the reference's hot-path routines over the C ABI.

Method names follow the reference Fortran routines they replace, with the
same argument meaning; arrays use the reference layouts ((NNODE,3) node
fields, 0-based indices on the Python side).  All compute runs in libofd.so
on the GPU; this module only moves arrays and sequences calls.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

# nonzero arrays of the coupled NS + VOF system and their entries per nonzero
NSVOF_BLOCKS = ("k11", "g", "d1", "m", "ls", "lsu", "pls", "uls")
_NSVOF_WIDTH = dict(k11=9, g=3, d1=3, m=1, ls=1, lsu=3, pls=1, uls=3)

# A cycle of the restarted solves stops once its preconditioned residual has
# fallen by this factor (or after kspace iterations).  It is the same for
# both precisions: a looser one ends cycles early and throws their Krylov
# space away at each restart (on the 16^3 cavity's first Newton system,
# 1e-2 takes 71 GMRES(30) cycles to a 1e-8 true residual, 1e-4 takes 16).
INNER_TOL = 1e-4


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _soa(a):
    """(N,3) -> column-major (the reference's (NNODE,3) memory order)."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)


class Context:
    """One device context (module state of aAdjKeep/commonvars, on the GPU)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = ctypes.c_void_p()
        L.check(self.lib.ofd_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.nnode = self.nelem = self.nshl = 0
        self.icnt = 0

    def close(self):
        if self.h:
            self.lib.ofd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- setup (input_mesh / genSparStruc / getparam / setBCs) ----
    def upload_mesh(self, mesh):
        ien = np.ascontiguousarray((np.asarray(mesh.ien, np.int32) + 1).T)
        elm = np.ascontiguousarray(mesh.elm_id, dtype=np.int32)
        xg = _soa(mesh.xg)
        self.nnode, self.nelem, self.nshl = mesh.nnode, mesh.nelem, mesh.nshl
        L.check(self.lib.ofd_upload_mesh(self.h, self.nnode, self.nelem, self.nshl,
                                         _ptr(ien), _ptr(elm), _ptr(xg)))

    def genSparStruc(self):
        """genSparStruc.f90:4-47 -> (col, row) 0-based."""
        n = ctypes.c_int64()
        L.check(self.lib.ofd_build_pattern(self.h, ctypes.byref(n)))
        self.icnt = int(n.value)
        col = np.zeros(self.nnode + 1, np.int32)
        row = np.zeros(self.icnt, np.int32)
        L.check(self.lib.ofd_get_pattern(self.h, _ptr(col), _ptr(row)))
        return col - 1, row - 1

    def check_pattern(self, col, row):
        """ofd_check_pattern with a 0-based (col, row): raises OfdError
        (OFD_ERR_ARG) unless it is the context's pattern."""
        c = np.ascontiguousarray(np.asarray(col, np.int32) + 1)
        r = np.ascontiguousarray(np.asarray(row, np.int32) + 1)
        L.check(self.lib.ofd_check_pattern(self.h, int(c.size) - 1, int(r.size), _ptr(c),
                                           _ptr(r)))

    def build_pattern(self):
        n = ctypes.c_int64()
        L.check(self.lib.ofd_build_pattern(self.h, ctypes.byref(n)))
        self.icnt = int(n.value)
        return self.icnt

    def set_params(self, p):
        """p: dict with the keys of _lib.PARAMS (grav may be given as 'grav')."""
        q = dict(p)
        if "grav" in q:
            q["gx"], q["gy"], q["gz"] = q.pop("grav")
        arr = np.array([q.get(k, 0.0) for k in L.PARAMS], dtype=np.float64)
        L.check(self.lib.ofd_set_params(self.h, _ptr(arr)))

    def set_subdomains(self, id1, id2):
        L.check(self.lib.ofd_set_subdomains(self.h, int(id1), int(id2)))

    def set_ibc(self, ibc):
        """IBC (NNODE, 4..8) int: columns u,v,w,p[,.,T,..] (column 6 = T)."""
        b = np.asarray(ibc, np.int32)
        nc = min(b.shape[1], 8)
        a = np.ascontiguousarray(b[:, :nc].T)
        L.check(self.lib.ofd_set_ibc(self.h, _ptr(a), nc))
        self._ibc_bc = (a, nc)

    def reset_ibc(self):
        """IBC back to the flags last given to set_ibc, dropping the fringe
        receivers flagged since: IBC(:,:) = 0 then setBCs_CFD, as each
        setBCs_overset call starts (solveOverset.f90:33-34, setBCs.f90:134)."""
        if getattr(self, "_ibc_bc", None) is None:
            raise RuntimeError("reset_ibc before set_ibc")
        a, nc = self._ibc_bc
        L.check(self.lib.ofd_set_ibc(self.h, _ptr(a), nc))

    def set_weak_faces(self, f2e, face_or):
        """Faces of the weak-Dirichlet boundaries, bound(b)%F2E and
        bound(b)%FACE_OR (input.f90:80-92), here with 0-based elements; hex
        FACE_OR 1..6 = xi -1/+1, eta -1/+1, zeta -1/+1.  assembleNavSto then
        adds FaceAssembly_NS_weak and FaceAssembly_NS_outflow over them."""
        e = np.ascontiguousarray(np.asarray(f2e, np.int32) + 1)
        o = np.ascontiguousarray(face_or, dtype=np.int32)
        if e.shape != o.shape:
            raise ValueError("f2e and face_or differ in length")
        L.check(self.lib.ofd_set_weak_faces(self.h, int(e.size), _ptr(e), _ptr(o)))

    def set_bcvals(self, vals):
        a = _soa(vals)
        L.check(self.lib.ofd_set_bcvals(self.h, _ptr(a)))

    def apply_dirichlet(self):
        """setBCs_CFD value part (setBCs.f90:84-91) on the device."""
        L.check(self.lib.ofd_apply_dirichlet(self.h))

    def set_field(self, name, arr):
        a = _f64(arr).ravel() if np.ndim(arr) == 1 else _soa(arr)
        L.check(self.lib.ofd_set_field(self.h, L.FIELD[name.upper()], _ptr(a)))

    def get_field(self, name):
        name = name.upper()
        if name in L.SCALAR_FIELDS:
            out = np.zeros(self.nnode)
            L.check(self.lib.ofd_get_field(self.h, L.FIELD[name], _ptr(out)))
            return out
        out = np.zeros((3, self.nnode))
        L.check(self.lib.ofd_get_field(self.h, L.FIELD[name], _ptr(out)))
        return out.T.copy()

    def set_state(self, fields: dict):
        for k, v in fields.items():
            self.set_field(k, v)

    # ---- assembly (assembleNavSto.f90:4-72) ----
    def assembleNavSto(self, assemble_tensor_flag=6, subdomain_flag=1):
        cfl = ctypes.c_double()
        L.check(self.lib.ofd_assemble_ns(self.h, int(assemble_tensor_flag),
                                         int(subdomain_flag), ctypes.byref(cfl)))
        return cfl.value

    def get_rhs(self):
        u = np.zeros((3, self.nnode))
        p = np.zeros(self.nnode)
        L.check(self.lib.ofd_get_rhs(self.h, _ptr(u), _ptr(p)))
        return u.T.copy(), p

    def set_rhs(self, rhsu, rhsp):
        u = _soa(rhsu)
        p = _f64(rhsp)
        L.check(self.lib.ofd_set_rhs(self.h, _ptr(u), _ptr(p)))

    def get_lhs(self):
        k = np.zeros((self.icnt, 9))
        g = np.zeros((self.icnt, 3))
        d = np.zeros((self.icnt, 3))
        m = np.zeros(self.icnt)
        L.check(self.lib.ofd_get_lhs(self.h, _ptr(k), _ptr(g), _ptr(d), _ptr(m)))
        return dict(k11=k, g=g, d1=d, m=m)

    def set_lhs(self, A):
        k, g, d, m = (_f64(A[n]) for n in ("k11", "g", "d1", "m"))
        L.check(self.lib.ofd_set_lhs(self.h, _ptr(k), _ptr(g), _ptr(d), _ptr(m)))

    # ---- linear algebra (SparseProd.f90, SparseGMRES_NS.f90) ----
    def SparseProdUP_3D(self, xu, xp, precision=L.OFD_FP64):
        xu_, xp_ = _soa(xu), _f64(xp)
        yu = np.zeros((3, self.nnode))
        yp = np.zeros(self.nnode)
        L.check(self.lib.ofd_spmv_up(self.h, precision, _ptr(xu_), _ptr(xp_), _ptr(yu),
                                     _ptr(yp)))
        return yu.T.copy(), yp

    def SparseGMRES_up(self, tol, kspace, kspace_min=0, precision=L.OFD_FP64,
                       ortho=L.OFD_ORTHO_MGS):
        su = np.zeros((3, self.nnode))
        sp = np.zeros(self.nnode)
        it = ctypes.c_int()
        res = ctypes.c_double()
        L.check(self.lib.ofd_gmres_up(self.h, precision, ortho, float(tol), int(kspace),
                                      int(kspace_min), _ptr(su), _ptr(sp),
                                      ctypes.byref(it), ctypes.byref(res)))
        return dict(solu=su.T.copy(), solp=sp, iters=it.value, resid=res.value)

    def solve(self, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS2,
              inner_tol=None, kspace=30, max_cycles=50, raise_noconv=True):
        if inner_tol is None:
            inner_tol = INNER_TOL
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.lib.ofd_solve(self.h, precision, ortho, float(rtol), float(inner_tol),
                                int(kspace), int(max_cycles), ctypes.byref(it),
                                ctypes.byref(cyc), ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != 5):
            L.check(rc)
        return dict(iters=it.value, cycles=cyc.value, relres=rel.value,
                    converged=rc == L.OFD_OK)

    def solve_history(self, maxn=4096):
        buf = (ctypes.c_double * maxn)()
        n = ctypes.c_int()
        L.check(self.lib.ofd_solve_history(self.h, buf, maxn, ctypes.byref(n)))
        return [buf[i] for i in range(min(n.value, maxn))]

    # ---- time / Newton drivers ----
    def newton_update(self):
        L.check(self.lib.ofd_newton_update(self.h))

    def predictor(self):
        L.check(self.lib.ofd_predictor(self.h))

    def advance(self):
        L.check(self.lib.ofd_advance(self.h))

    def rhs_norms(self):
        out = (ctypes.c_double * 2)()
        L.check(self.lib.ofd_rhs_norms(self.h, out))
        return float(out[0]), float(out[1])

    # ---- temperature equation of the quenching solver (solveQuenching.f90) ----
    def set_quench_params(self, q):
        """q: dict with the keys of _lib.QPARAMS (fluid_id defaults to 102,
        lsc_kdc, the VOF discontinuity-capturing constant, to 0)."""
        d = dict(q)
        d.setdefault("fluid_id", 102)
        d.setdefault("lsc_kdc", 0.0)
        arr = np.array([float(d[k]) for k in L.QPARAMS], dtype=np.float64)
        L.check(self.lib.ofd_set_quench_params(self.h, _ptr(arr)))

    def assemble_tem(self, tensor_flag=6):
        """IntElmAss_Tem_Quenching_STAB + BCLhs_tem + FillSparseMat_tem."""
        L.check(self.lib.ofd_assemble_tem(self.h, int(tensor_flag)))

    def get_tem_system(self):
        lhs = np.zeros(self.icnt)
        rhs = np.zeros(self.nnode)
        L.check(self.lib.ofd_get_tem_system(self.h, _ptr(lhs), _ptr(rhs)))
        return lhs, rhs

    def set_tem_system(self, lhs=None, rhs=None):
        lp = None if lhs is None else _f64(lhs)
        rp = None if rhs is None else _f64(rhs)
        L.check(self.lib.ofd_set_tem_system(self.h, None if lp is None else _ptr(lp),
                                            None if rp is None else _ptr(rp)))

    def SparseProd_LS(self, x, precision=L.OFD_FP64):
        x_ = _f64(x)
        y = np.zeros(self.nnode)
        L.check(self.lib.ofd_spmv_tem(self.h, precision, _ptr(x_), _ptr(y)))
        return y

    def SparseGMRES_tem(self, tol, kspace, kspace_min=0, precision=L.OFD_FP64,
                        ortho=L.OFD_ORTHO_MGS):
        sol = np.zeros(self.nnode)
        it = ctypes.c_int()
        res = ctypes.c_double()
        L.check(self.lib.ofd_gmres_tem(self.h, precision, ortho, float(tol), int(kspace),
                                       int(kspace_min), _ptr(sol), ctypes.byref(it),
                                       ctypes.byref(res)))
        return dict(sol=sol, iters=it.value, resid=res.value)

    def solve_tem(self, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                  inner_tol=None, kspace=30, max_cycles=50, raise_noconv=True):
        if inner_tol is None:
            inner_tol = INNER_TOL
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.lib.ofd_solve_tem(self.h, precision, ortho, float(rtol), float(inner_tol),
                                    int(kspace), int(max_cycles), ctypes.byref(it),
                                    ctypes.byref(cyc), ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != 5):
            L.check(rc)
        return dict(iters=it.value, cycles=cyc.value, relres=rel.value,
                    converged=rc == L.OFD_OK)

    def tem_update(self):
        L.check(self.lib.ofd_tem_update(self.h))

    def tem_rhs_norm(self):
        out = ctypes.c_double()
        L.check(self.lib.ofd_tem_rhs_norm(self.h, ctypes.byref(out)))
        return out.value

    # ---- coupled NS + VOF system of the quenching solver ----
    def assemble_nsvof(self, tensor_flag=6):
        """IntElmAss_NSVOF_Quenching_STAB (solveQuenching.f90) with the
        BCLhs_NS_conv / FillSparseMat_NS_conv scatter; returns the CFL maximum."""
        cfl = ctypes.c_double()
        L.check(self.lib.ofd_assemble_nsvof(self.h, int(tensor_flag), ctypes.byref(cfl)))
        return cfl.value

    def get_nsvof_system(self):
        """The eight nonzero arrays in the reference's layout (LHSK11 (icnt,9),
        LHSG, LHSD1, LHSLSU, LHSULS (icnt,3), LHSM, LHSLS, LHSPLS (icnt)) and
        RHSGu (N,3), RHSGp, RHSGls (N)."""
        Z, N = self.icnt, self.nnode
        A = {k: np.zeros((Z, w) if w > 1 else Z) for k, w in _NSVOF_WIDTH.items()}
        ru = np.zeros((3, N))
        rp = np.zeros(N)
        rl = np.zeros(N)
        L.check(self.lib.ofd_get_nsvof_system(self.h, *[_ptr(A[k]) for k in NSVOF_BLOCKS],
                                              _ptr(ru), _ptr(rp), _ptr(rl)))
        return A, dict(rhsu=ru.T.copy(), rhsp=rp, rhsls=rl)

    def set_nsvof_system(self, A=None, rhs=None):
        mats = [None] * 8 if A is None else [_f64(A[k]) for k in NSVOF_BLOCKS]
        rh = [None] * 3 if rhs is None else [_soa(rhs["rhsu"]), _f64(rhs["rhsp"]),
                                             _f64(rhs["rhsls"])]
        L.check(self.lib.ofd_set_nsvof_system(
            self.h, *[None if a is None else _ptr(a) for a in mats + rh]))

    def SparseProd_NS_conv(self, xu, xp, xl, precision=L.OFD_FP64):
        xu_, xp_, xl_ = _soa(xu), _f64(xp), _f64(xl)
        N = self.nnode
        yu = np.zeros((3, N))
        yp = np.zeros(N)
        yl = np.zeros(N)
        L.check(self.lib.ofd_spmv_nsvof(self.h, precision, _ptr(xu_), _ptr(xp_), _ptr(xl_),
                                        _ptr(yu), _ptr(yp), _ptr(yl)))
        return yu.T.copy(), yp, yl

    def SparseGMRES_ns_conv(self, tol, kspace, kspace_min=0, precision=L.OFD_FP64,
                            ortho=L.OFD_ORTHO_MGS):
        N = self.nnode
        su = np.zeros((3, N))
        sp = np.zeros(N)
        sl = np.zeros(N)
        it = ctypes.c_int()
        res = ctypes.c_double()
        L.check(self.lib.ofd_gmres_nsvof(self.h, precision, ortho, float(tol), int(kspace),
                                         int(kspace_min), _ptr(su), _ptr(sp), _ptr(sl),
                                         ctypes.byref(it), ctypes.byref(res)))
        return dict(solu=su.T.copy(), solp=sp, solls=sl, iters=it.value, resid=res.value)

    def solve_nsvof(self, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                    inner_tol=None, kspace=30, max_cycles=50, raise_noconv=True):
        """Restarted SparseGMRES_ns_conv with an FP64 outer residual.  ``ortho``
        may carry L.OFD_PREC_RIGHT (with CGS1 or CGS2): the block-Jacobi
        preconditioner on the right, so each cycle minimises the true
        residual."""
        if inner_tol is None:
            inner_tol = INNER_TOL
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.lib.ofd_solve_nsvof(self.h, precision, ortho, float(rtol), float(inner_tol),
                                      int(kspace), int(max_cycles), ctypes.byref(it),
                                      ctypes.byref(cyc), ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != 5):
            L.check(rc)
        return dict(iters=it.value, cycles=cyc.value, relres=rel.value,
                    converged=rc == L.OFD_OK)

    def nsvof_update(self):
        L.check(self.lib.ofd_nsvof_update(self.h))

    def nsvof_rhs_norms(self):
        out = (ctypes.c_double * 3)()
        L.check(self.lib.ofd_nsvof_rhs_norms(self.h, out))
        return tuple(float(v) for v in out)

    # ---- overset fringe (interpolation.f90:5-76, setBCs.f90:124-167) ----
    def interpolation(self, src_id, xr, dr=None):
        """Donor search and interpolation of ug at receivers xr + dr
        ((nrecv,3) each).  Returns (donor, val): donor 1-based element
        index (0: none), val (nrecv,3)."""
        xr_ = _soa(xr)
        nr = xr_.shape[1]
        dr_ = _soa(np.zeros((nr, 3)) if dr is None else dr)
        donor = np.zeros(nr, np.int32)
        val = np.zeros((3, nr))
        L.check(self.lib.ofd_fringe_interp(self.h, int(src_id), nr, _ptr(xr_), _ptr(dr_),
                                           _ptr(donor), _ptr(val)))
        return donor, val.T.copy()

    def apply_fringe(self, nodes, val):
        """setBCs_overset step 3 for 1-based ``nodes`` with values (n,3)."""
        nodes_ = np.ascontiguousarray(nodes, dtype=np.int32)
        val_ = _soa(val)
        L.check(self.lib.ofd_apply_fringe(self.h, int(nodes_.size), _ptr(nodes_), _ptr(val_)))

    # ---- adjoint of the NS Newton step (no reference code) ----
    def SparseProdUP_3D_T(self, xu, xp, precision=L.OFD_FP64):
        """y = LHS^T x for the assembled LHS."""
        xu_, xp_ = _soa(xu), _f64(xp)
        yu = np.zeros((3, self.nnode))
        yp = np.zeros(self.nnode)
        L.check(self.lib.ofd_spmv_up_t(self.h, precision, _ptr(xu_), _ptr(xp_), _ptr(yu),
                                       _ptr(yp)))
        return yu.T.copy(), yp

    def _adj_solve(self, fn, args, gu, gp, raise_noconv):
        gu_, gp_ = _soa(gu), _f64(gp)
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = fn(self.h, *args, _ptr(gu_), _ptr(gp_), ctypes.byref(it), ctypes.byref(cyc),
                ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != 5):
            L.check(rc)
        return dict(lamu=self.get_field("LAMU"), lamp=self.get_field("LAMP"),
                    iters=it.value, cycles=cyc.value, relres=rel.value,
                    converged=rc == L.OFD_OK)

    def solve_t(self, gu, gp, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                inner_tol=None, kspace=30, max_cycles=100, raise_noconv=True):
        """LHS^T lambda = g for the assembled LHS."""
        if inner_tol is None:
            inner_tol = INNER_TOL
        return self._adj_solve(self.lib.ofd_solve_t,
                               (precision, ortho, float(rtol), float(inner_tol), int(kspace),
                                int(max_cycles)), gu, gp, raise_noconv)

    def vjp_ns(self, lu, lp, subdomain_flag=1):
        """Gradients of lambda . F (the NS residual as assembled, before the
        Dirichlet projection) with respect to the alpha-stage u, ac, p and
        the per-node viscosity scale."""
        lu_, lp_ = _soa(lu), _f64(lp)
        N = self.nnode
        gu = np.zeros((3, N))
        gac = np.zeros((3, N))
        gp = np.zeros(N)
        gm = np.zeros(N)
        L.check(self.lib.ofd_vjp_ns(self.h, int(subdomain_flag), _ptr(lu_), _ptr(lp_),
                                    _ptr(gu), _ptr(gac), _ptr(gp), _ptr(gm)))
        return dict(u=gu.T.copy(), ac=gac.T.copy(), p=gp, mus=gm)

    def jt_apply(self, lu, lp, subdomain_flag=1):
        """y = J_bc^T lambda, the exact Jacobian of the residual with respect
        to the Newton increments with the Dirichlet projection, transposed."""
        lu_, lp_ = _soa(lu), _f64(lp)
        yu = np.zeros((3, self.nnode))
        yp = np.zeros(self.nnode)
        L.check(self.lib.ofd_jt_apply(self.h, int(subdomain_flag), _ptr(lu_), _ptr(lp_),
                                      _ptr(yu), _ptr(yp)))
        return yu.T.copy(), yp

    def set_residual_form(self, form):
        """0: the reference's NS residual (uprime = 0); 1: with the fine-scale
        velocity uprime = -tauM rLi."""
        L.check(self.lib.ofd_set_residual_form(self.h, int(form)))

    def j_apply(self, vu, vp, subdomain_flag=1):
        """y = J_bc v for a Newton increment v (forward tangent)."""
        vu_, vp_ = _soa(vu), _f64(vp)
        yu = np.zeros((3, self.nnode))
        yp = np.zeros(self.nnode)
        L.check(self.lib.ofd_j_apply(self.h, int(subdomain_flag), _ptr(vu_), _ptr(vp_),
                                     _ptr(yu), _ptr(yp)))
        return yu.T.copy(), yp

    def solve_newton(self, rtol=1e-10, subdomain_flag=1, precision=L.OFD_FP64,
                     ortho=L.OFD_ORTHO_CGS1, inner_tol=1e-3, kspace=30, max_cycles=20,
                     raise_noconv=True):
        """The consistent Newton correction J_bc d = RHS (flexible GMRES
        preconditioned by the assembled LHS); apply it with newton_update."""
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.lib.ofd_solve_newton(self.h, int(subdomain_flag), precision, ortho,
                                       float(rtol), float(inner_tol), int(kspace),
                                       int(max_cycles), ctypes.byref(it), ctypes.byref(cyc),
                                       ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != 5):
            L.check(rc)
        return dict(iters=it.value, cycles=cyc.value, relres=rel.value,
                    converged=rc == L.OFD_OK)

    def solve_adjoint(self, gu, gp, rtol=1e-8, subdomain_flag=1, precision=L.OFD_FP64,
                      ortho=L.OFD_ORTHO_CGS1, inner_tol=1e-3, kspace=30, max_cycles=20,
                      raise_noconv=True):
        """J_bc^T lambda = g: flexible GMRES on the exact operator (element
        VJP), right-preconditioned by LHS^T solves."""
        return self._adj_solve(self.lib.ofd_solve_adjoint,
                               (int(subdomain_flag), precision, ortho, float(rtol),
                                float(inner_tol), int(kspace), int(max_cycles)), gu, gp,
                               raise_noconv)

    # ---- adjoint of the quenching temperature equation (no reference code) ----
    def SparseProd_LS_T(self, x, precision=L.OFD_FP64):
        """y = LHSTem^T x."""
        x_ = _f64(x)
        y = np.zeros(self.nnode)
        L.check(self.lib.ofd_spmv_tem_t(self.h, precision, _ptr(x_), _ptr(y)))
        return y

    def _tem_adj_solve(self, fn, args, g, raise_noconv):
        g_ = _f64(g)
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = fn(self.h, *args, _ptr(g_), ctypes.byref(it), ctypes.byref(cyc),
                ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != 5):
            L.check(rc)
        return dict(lam=self.get_field("LAMT"), iters=it.value, cycles=cyc.value,
                    relres=rel.value, converged=rc == L.OFD_OK)

    def solve_tem_t(self, g, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                    inner_tol=None, kspace=30, max_cycles=100, raise_noconv=True):
        """LHSTem^T lambda = g for the assembled LHSTem."""
        if inner_tol is None:
            inner_tol = INNER_TOL
        return self._tem_adj_solve(self.lib.ofd_solve_tem_t,
                                   (precision, ortho, float(rtol), float(inner_tol),
                                    int(kspace), int(max_cycles)), g, raise_noconv)

    def vjp_tem(self, lam):
        """Gradients of lambda . R_T (the temperature residual as assembled,
        before the Dirichlet projection) with respect to the alpha-stage T,
        rT, u and the per-node conductivity scale."""
        lam_ = _f64(lam)
        N = self.nnode
        gT, grT, gks = np.zeros(N), np.zeros(N), np.zeros(N)
        gu = np.zeros((3, N))
        L.check(self.lib.ofd_vjp_tem(self.h, _ptr(lam_), _ptr(gT), _ptr(grT), _ptr(gu),
                                     _ptr(gks)))
        return dict(T=gT, rT=grT, u=gu.T.copy(), ks=gks)

    def jt_apply_tem(self, lam):
        """y = J_T,bc^T lambda (exact Jacobian of the temperature residual with
        respect to the Newton increments, with the Dirichlet projection)."""
        lam_ = _f64(lam)
        y = np.zeros(self.nnode)
        L.check(self.lib.ofd_jt_apply_tem(self.h, _ptr(lam_), _ptr(y)))
        return y

    def solve_tem_adjoint(self, g, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS1,
                          inner_tol=1e-3, kspace=30, max_cycles=20, raise_noconv=True):
        """J_T,bc^T lambda = g: flexible GMRES on the exact operator,
        right-preconditioned by LHSTem^T solves."""
        return self._tem_adj_solve(self.lib.ofd_solve_tem_adjoint,
                                   (precision, ortho, float(rtol), float(inner_tol),
                                    int(kspace), int(max_cycles)), g, raise_noconv)

    # ---- exact NS + VOF Jacobian and the coupled quench adjoint (no reference code) ----
    def assemble_nsvof_jac(self):
        """Replace the stored NS + VOF matrix by J_bc, the exact Jacobian of
        the assembled RHS with respect to the Newton increments (dual-number
        element kernels), rows and columns of the constrained dofs cleared;
        solve_nsvof then takes exact Newton steps."""
        L.check(self.lib.ofd_assemble_nsvof_jac(self.h))

    def vjp_nsvof(self, lu, lp, lphi):
        """Gradients of lambda . R (the NS + VOF residual as assembled, before
        the Dirichlet projection) with respect to the alpha-stage u, ac, p,
        phi, rphi and T."""
        N = self.nnode
        lu_, lp_, ll_ = _soa(lu), _f64(lp), _f64(lphi)
        g = np.zeros(10 * N)
        L.check(self.lib.ofd_vjp_nsvof(self.h, _ptr(lu_), _ptr(lp_), _ptr(ll_), _ptr(g)))
        g = g.reshape(10, N)
        return dict(u=g[0:3].T.copy(), ac=g[3:6].T.copy(), p=g[6].copy(), phi=g[7].copy(),
                    rphi=g[8].copy(), T=g[9].copy())

    def vjp_tem_dual(self, lam):
        """Gradients of lambda . R_T with respect to the alpha-stage T, rT, u,
        phi and the conductivity scale (dual-number element kernel)."""
        N = self.nnode
        lam_ = _f64(lam)
        g = np.zeros(7 * N)
        L.check(self.lib.ofd_vjp_tem_dual(self.h, _ptr(lam_), _ptr(g)))
        g = g.reshape(7, N)
        return dict(T=g[0].copy(), rT=g[1].copy(), u=g[2:5].T.copy(), phi=g[5].copy(),
                    ks=g[6].copy())

    def SparseProd_NS_conv_T(self, xu, xp, xl):
        """y = A^T x for the stored NS + VOF matrix (FP64)."""
        N = self.nnode
        xu_, xp_, xl_ = _soa(xu), _f64(xp), _f64(xl)
        yu, yp, yl = np.zeros(3 * N), np.zeros(N), np.zeros(N)
        L.check(self.lib.ofd_spmv_nsvof_t(self.h, _ptr(xu_), _ptr(xp_), _ptr(xl_), _ptr(yu),
                                          _ptr(yp), _ptr(yl)))
        return yu.reshape(3, N).T.copy(), yp, yl

    def solve_nsvof_t(self, gu, gp, gl, rtol=1e-10, precision=L.OFD_FP64,
                      ortho=L.OFD_ORTHO_CGS1 | L.OFD_PREC_RIGHT, inner_tol=1e-3, kspace=30,
                      max_cycles=20, raise_noconv=True):
        """A^T lambda = g for the stored NS + VOF matrix: flexible GMRES on the
        exact transposed product, right-preconditioned by restarted GMRES on
        the stored transpose (its (phi, p) coupling dropped) to inner_tol."""
        N = self.nnode
        gu_, gp_, gl_ = _soa(gu), _f64(gp), _f64(gl)
        lu, lp, ll = np.zeros(3 * N), np.zeros(N), np.zeros(N)
        it = ctypes.c_int()
        cyc = ctypes.c_int()
        rel = ctypes.c_double()
        rc = self.lib.ofd_solve_nsvof_t(self.h, precision, ortho, float(rtol), float(inner_tol),
                                        int(kspace), int(max_cycles), _ptr(gu_), _ptr(gp_),
                                        _ptr(gl_), _ptr(lu), _ptr(lp), _ptr(ll),
                                        ctypes.byref(it), ctypes.byref(cyc), ctypes.byref(rel))
        if rc != L.OFD_OK and (raise_noconv or rc != L.OFD_ERR_NOCONV):
            L.check(rc)
        return dict(lamu=lu.reshape(3, N).T.copy(), lamp=lp, lamphi=ll, iters=it.value,
                    cycles=cyc.value, relres=rel.value, converged=rc == L.OFD_OK)

    # ---- measurement ----
    def time_quench(self, what, arg, reps=5):
        """Mean device ms of one quenching-solver kernel (L.OFD_TIME_*; arg is
        the tensor flag for an assembly, the precision for a product)."""
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_quench(self.h, int(what), int(arg), int(reps),
                                         ctypes.byref(ms)))
        return ms.value

    def time_adjoint(self, what, reps=5):
        """Mean device ms of one adjoint kernel (L.OFD_TIME_VJP_NS, ..._JT_NS,
        ..._VJP_TEM, ..._JT_TEM) on the current state and the last uploaded
        lambda."""
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_adjoint(self.h, int(what), int(reps), ctypes.byref(ms)))
        return ms.value

    def time_spmv(self, precision=L.OFD_FP64, reps=20):
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_spmv(self.h, precision, int(reps), ctypes.byref(ms)))
        return ms.value

    def time_assemble(self, tensor_flag=6, reps=5):
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_assemble(self.h, int(tensor_flag), int(reps),
                                           ctypes.byref(ms)))
        return ms.value

    def time_read(self, nbytes=1 << 30, reps=10):
        """Average ms of one read-only stream of ``nbytes``."""
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_read(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def time_copy(self, nbytes=1 << 28, reps=10):
        """Average ms of one device copy of ``nbytes`` (reads and writes
        ``nbytes`` each)."""
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_copy(self.h, int(nbytes), int(reps), ctypes.byref(ms)))
        return ms.value

    def time_faces(self, tensor_flag=6, reps=20):
        """Mean ms of the weak-face launches of one assembly (HIP events);
        the system must be assembled again afterwards."""
        ms = ctypes.c_double()
        L.check(self.lib.ofd_time_faces(self.h, int(tensor_flag), int(reps), ctypes.byref(ms)))
        return ms.value

    def probe_set(self, every):
        """Sample every ``every``-th Arnoldi product launch of later solves
        with HIP timing events (0: off)."""
        L.check(self.lib.ofd_probe_set(self.h, int(every)))

    def probe_read(self):
        """(launches sampled, summed ms) since the last read; resets both."""
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        L.check(self.lib.ofd_probe_read(self.h, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def device_bytes(self):
        b = ctypes.c_int64()
        L.check(self.lib.ofd_ctx_bytes(self.h, ctypes.byref(b)))
        return b.value


def setup(prob, device=0) -> Context:
    """Context holding a configs.Problem: mesh, pattern, parameters, Dirichlet
    flags and values, initial fields."""
    ctx = Context(device)
    ctx.upload_mesh(prob.mesh)
    ctx.build_pattern()
    ctx.set_params(prob.params)
    ctx.set_ibc(prob.ibc)
    ctx.set_bcvals(prob.bcval)
    ctx.set_state(prob.fields)
    faces = getattr(prob, "weak_faces", None)
    if faces is not None:
        ctx.set_weak_faces(*faces)
    return ctx


def begin_step(ctx: Context):
    """Predictor (driver.f90:83-85) then Dirichlet values (setBCs_CFD)."""
    ctx.predictor()
    ctx.apply_dirichlet()


def newton_step(ctx: Context, rtol=1e-8, precision=L.OFD_FP64, ortho=L.OFD_ORTHO_CGS2,
                kspace=30, inner_tol=None, max_cycles=50, raise_noconv=True,
                subdomain_flag=1):
    """One Newton iteration of solveOverset_stag (solveOverset.f90:55-121) on
    the elements of ``subdomain_flag``: assemble VEC+MAT, solve to rtol,
    update, re-assemble VEC and return the residual norms."""
    ctx.assembleNavSto(6, subdomain_flag)
    info = ctx.solve(rtol=rtol, precision=precision, ortho=ortho, kspace=kspace,
                     inner_tol=inner_tol, max_cycles=max_cycles,
                     raise_noconv=raise_noconv)
    ctx.newton_update()
    ctx.assembleNavSto(2, subdomain_flag)
    info["res"] = ctx.rhs_norms()
    return info


def quench_step(ctx: Context, nnewton=3, rtol=1e-8, precision=L.OFD_FP64,
                 ortho=L.OFD_ORTHO_CGS1, kspace=30, max_cycles=200, raise_noconv=True,
                 ntol=None, nsvof_right=False):
    """One time step of the quenching solver, solmultiphasethermofluid_stag
    (solveFlow.f90:4-164) after the predictor (driver.f90:83-85): per
    Newton iteration the NS + VOF system (assembleQuenching VEC+MAT with the
    NS and VOF fields, restarted SparseGMRES_ns_conv to rtol, the update of
    solveFlow.f90:99-104), then the temperature system (VEC+MAT with the
    temperature field, restarted SparseGMRES_tem, the update of :139-140),
    then the residuals of both (:145-155).  With ``ntol`` = (u, p, phi, T)
    relative tolerances the loop stops once every residual has fallen by its
    tolerance from the first (check_convergence, solveFlow.f90:178-194);
    without it, ``nnewton`` iterations run.  ``nsvof_right`` applies the
    NS + VOF system's block-Jacobi preconditioner on the right
    (L.OFD_PREC_RIGHT).  Returns the per-iteration solver information, with
    the residual norms (u, p, phi, T) after each iteration (``res``) and
    after the predictor (``res0``)."""
    ctx.predictor()
    ctx.apply_dirichlet()
    ctx.assemble_nsvof(2)
    ctx.assemble_tem(2)
    r0 = (*ctx.nsvof_rhs_norms(), ctx.tem_rhs_norm())
    out = []
    for _ in range(nnewton):
        ctx.assemble_nsvof(6)
        iv = ctx.solve_nsvof(rtol=rtol, precision=precision,
                             ortho=ortho | (L.OFD_PREC_RIGHT if nsvof_right else 0),
                             kspace=kspace, max_cycles=max_cycles, raise_noconv=raise_noconv)
        ctx.nsvof_update()
        ctx.assemble_tem(6)
        it = ctx.solve_tem(rtol=rtol, precision=precision, ortho=ortho, kspace=kspace,
                           max_cycles=max_cycles, raise_noconv=raise_noconv)
        ctx.tem_update()
        ctx.assemble_nsvof(2)
        ctx.assemble_tem(2)
        res = (*ctx.nsvof_rhs_norms(), ctx.tem_rhs_norm())
        out.append(dict(nsvof=iv, tem=it, res=res, res0=r0))
        if ntol is not None and all(r <= t * max(1e-15, r0_) for r, t, r0_ in
                                    zip(res, ntol, r0)):
            break
    return out


def fill_fringe(ctx: Context, fringe, xg):
    """setBCs_overset's fringe step (setBCs.f90:124-167) for every block:
    donor search and interpolation of ug in the source block
    (interpolation_tet, interpolation.f90:5-76, or the hex extension), then
    IBC(k,1:3) = 1, ug = the interpolated value and the consistent acg at
    the receivers.  ``fringe``: {name: (0-based receiver nodes, source
    ELM_ID)}.  Returns the donors per block."""
    donors = {}
    for name, (nodes, src) in fringe.items():
        donor, val = ctx.interpolation(src, xg[nodes])
        if np.any(donor <= 0):
            raise RuntimeError(f"fringe {name}: {int(np.sum(donor <= 0))} receivers have no "
                               "donor")
        ctx.apply_fringe(np.asarray(nodes, np.int32) + 1, val)
        donors[name] = donor
    return donors


def overset_step(ctx: Context, fringe, xg, nnewton=3, rtol=1e-8, precision=L.OFD_FP64,
                 ortho=L.OFD_ORTHO_CGS1, kspace=30, inner_tol=None, max_cycles=400,
                 raise_noconv=True):
    """One time step of an overset pair coupled in one system: predictor
    and Dirichlet values, then per Newton iteration the fringe of every
    block from the other's latest iterate, one assembly launch over both
    blocks (subdomain flags 1 | 2) and one solve of the coupled system.
    The physical Dirichlet values go only on the physical flags: the
    receivers flagged by the previous step's fringe are cleared first, so
    they keep their interpolated velocity into the next donor search.
    Returns the per-iteration solver information."""
    ctx.reset_ibc()
    begin_step(ctx)
    out = []
    for _ in range(nnewton):
        fill_fringe(ctx, fringe, xg)
        out.append(newton_step(ctx, rtol=rtol, precision=precision, ortho=ortho,
                               kspace=kspace, inner_tol=inner_tol, max_cycles=max_cycles,
                               raise_noconv=raise_noconv, subdomain_flag=3))
    return out


# writeSol's fields (iosol.f90:400-429): the current values, except the
# temperature and its rate, which writeSol takes from the old arrays
_WRITESOL = (("dg", "DG"), ("ug", "UG"), ("ugm", "UGM"), ("acg", "ACG"), ("acgm", "ACGM"),
             ("phig", "PHIG"), ("rphig", "RPHIG"), ("pg", "PG"), ("Tg", "TGOLD"),
             ("rTg", "RTGOLD"))
# readSol's targets (iosol.f90:146-175): every field into the old arrays
_READSOL = (("dg", "DGOLD"), ("ug", "UGOLD"), ("ugm", "UGMOLD"), ("acg", "ACGOLD"),
            ("acgm", "ACGMOLD"), ("phig", "PHIGOLD"), ("rphig", "RPHIGOLD"), ("pg", "PGOLD"),
            ("Tg", "TGOLD"), ("rTg", "RTGOLD"))


def write_restart(ctx: Context, path, step, time):
    """writeSol (iosol.f90:377-454) from the context's fields."""
    from . import refio
    refio.write_restart(str(path), step, time, {k: ctx.get_field(f) for k, f in _WRITESOL})


def read_restart(ctx: Context, path):
    """readSol (iosol.f90:122-194) into the context's old fields; returns
    (step, time)."""
    from . import refio
    r = refio.read_restart(str(path), ctx.nnode)
    for k, f in _READSOL:
        ctx.set_field(f, r[k])
    return r["step"], r["time"]
