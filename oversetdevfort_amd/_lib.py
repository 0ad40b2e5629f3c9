"""ctypes binding of libofd.so (the C ABI declared in include/ofd.h).

The library is loaded from the package directory only; when it is missing
the import fails loudly -- there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libofd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ofd.h")

OFD_OK = 0
OFD_ERR_ARG = 1
OFD_ERR_HIP = 2
OFD_ERR_NEGDETJ = 3
OFD_ERR_NONFINITE = 4
OFD_ERR_NOCONV = 5
OFD_ERR_ALLOC = 6
OFD_ERR_UNSUPPORTED = 7
OFD_ASSEMBLE_VEC = 2
OFD_ASSEMBLE_MAT = 4
OFD_FP64 = 0
OFD_FP32 = 1
OFD_ORTHO_MGS = 0
OFD_ORTHO_CGS2 = 1
OFD_ORTHO_CGS1 = 2
# OR'ed into ortho of ofd_solve_nsvof: block Jacobi on the right
OFD_PREC_RIGHT = 16
OFD_TIME_ASM_NSVOF = 0
OFD_TIME_ASM_TEM = 1
OFD_TIME_SPMV_NSVOF = 2
OFD_TIME_SPMV_TEM = 3
OFD_TIME_VJP_NS = 0
OFD_TIME_JT_NS = 1
OFD_TIME_VJP_TEM = 2
OFD_TIME_JT_TEM = 3

FIELDS = ["XG", "DG", "DGOLD", "UG", "UGOLD", "UGM", "UGMOLD", "ACG", "ACGOLD",
          "ACGM", "ACGMOLD", "PG", "PGOLD", "SOLU", "SOLP",
          "PHIG", "PHIGOLD", "RPHIG", "RPHIGOLD", "TG", "TGOLD", "RTG", "RTGOLD", "SOLT",
          "SOLPHI", "LAMU", "LAMP", "MUSCALE", "KSCALE", "LAMT"]
FIELD = {n: i for i, n in enumerate(FIELDS)}
SCALAR_FIELDS = {"PG", "PGOLD", "SOLP", "PHIG", "PHIGOLD", "RPHIG", "RPHIGOLD", "TG",
                 "TGOLD", "RTG", "RTGOLD", "SOLT", "SOLPHI", "LAMP",
                 "MUSCALE", "KSCALE", "LAMT"}
PARAMS = ["almi", "alfi", "gami", "beti", "delt", "rho", "mu", "kdc", "gx", "gy",
          "gz"]
# material vector of ofd_set_quench_params
QPARAMS = ["rhoa", "rhow", "mua", "muw", "cpa", "cpw", "kappaa", "kappaw", "rhos", "mus",
           "cps", "kappas", "Ts", "lh", "c_cond", "c_evap", "c_dmdot", "tem_kdc", "tref",
           "fluid_id", "lsc_kdc"]

_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)

# argument types of every entry point of include/ofd.h
SIGNATURES = {
    "ofd_last_error": ([], ctypes.c_char_p),
    "ofd_version": ([], _c_int),
    "ofd_ctx_create": ([_c_int, ctypes.POINTER(_vp)], _c_int),
    "ofd_ctx_destroy": ([_vp], _c_int),
    "ofd_upload_mesh": ([_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp], _c_int),
    "ofd_build_pattern": ([_vp, _i64p], _c_int),
    "ofd_get_pattern": ([_vp, _vp, _vp], _c_int),
    "ofd_check_pattern": ([_vp, _c_int, ctypes.c_int64, _vp, _vp], _c_int),
    "ofd_set_params": ([_vp, _vp], _c_int),
    "ofd_set_subdomains": ([_vp, _c_int, _c_int], _c_int),
    "ofd_set_ibc": ([_vp, _vp, _c_int], _c_int),
    "ofd_set_bcvals": ([_vp, _vp], _c_int),
    "ofd_apply_dirichlet": ([_vp], _c_int),
    "ofd_set_field": ([_vp, _c_int, _vp], _c_int),
    "ofd_get_field": ([_vp, _c_int, _vp], _c_int),
    "ofd_assemble_ns": ([_vp, _c_int, _c_int, _dp], _c_int),
    "ofd_get_rhs": ([_vp, _vp, _vp], _c_int),
    "ofd_get_lhs": ([_vp, _vp, _vp, _vp, _vp], _c_int),
    "ofd_set_lhs": ([_vp, _vp, _vp, _vp, _vp], _c_int),
    "ofd_set_rhs": ([_vp, _vp, _vp], _c_int),
    "ofd_spmv_up": ([_vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ofd_gmres_up": ([_vp, _c_int, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp, _ip,
                      _dp], _c_int),
    "ofd_solve": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _ip, _ip,
                   _dp], _c_int),
    "ofd_solve_history": ([_vp, _dp, _c_int, _ip], _c_int),
    "ofd_newton_update": ([_vp], _c_int),
    "ofd_predictor": ([_vp], _c_int),
    "ofd_advance": ([_vp], _c_int),
    "ofd_rhs_norms": ([_vp, _dp], _c_int),
    "ofd_set_quench_params": ([_vp, _vp], _c_int),
    "ofd_assemble_tem": ([_vp, _c_int], _c_int),
    "ofd_get_tem_system": ([_vp, _vp, _vp], _c_int),
    "ofd_set_tem_system": ([_vp, _vp, _vp], _c_int),
    "ofd_spmv_tem": ([_vp, _c_int, _vp, _vp], _c_int),
    "ofd_gmres_tem": ([_vp, _c_int, _c_int, _c_dbl, _c_int, _c_int, _vp, _ip, _dp], _c_int),
    "ofd_solve_tem": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _ip, _ip, _dp],
                      _c_int),
    "ofd_tem_update": ([_vp], _c_int),
    "ofd_tem_rhs_norm": ([_vp, _dp], _c_int),
    "ofd_assemble_nsvof": ([_vp, _c_int, _dp], _c_int),
    "ofd_get_nsvof_system": ([_vp] + [_vp] * 11, _c_int),
    "ofd_set_nsvof_system": ([_vp] + [_vp] * 11, _c_int),
    "ofd_spmv_nsvof": ([_vp, _c_int] + [_vp] * 6, _c_int),
    "ofd_gmres_nsvof": ([_vp, _c_int, _c_int, _c_dbl, _c_int, _c_int, _vp, _vp, _vp, _ip,
                         _dp], _c_int),
    "ofd_solve_nsvof": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _ip, _ip,
                         _dp], _c_int),
    "ofd_nsvof_update": ([_vp], _c_int),
    "ofd_nsvof_rhs_norms": ([_vp, _vp], _c_int),
    "ofd_spmv_up_t": ([_vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ofd_solve_t": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _vp, _vp, _ip,
                     _ip, _dp], _c_int),
    "ofd_vjp_ns": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "ofd_jt_apply": ([_vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ofd_j_apply": ([_vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ofd_set_residual_form": ([_vp, _c_int], _c_int),
    "ofd_solve_newton": ([_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _ip,
                          _ip, _dp], _c_int),
    "ofd_solve_adjoint": ([_vp, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _vp,
                           _vp, _ip, _ip, _dp], _c_int),
    "ofd_spmv_tem_t": ([_vp, _c_int, _vp, _vp], _c_int),
    "ofd_solve_tem_t": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _vp, _ip, _ip,
                         _dp], _c_int),
    "ofd_vjp_tem": ([_vp] + [_vp] * 5, _c_int),
    "ofd_jt_apply_tem": ([_vp, _vp, _vp], _c_int),
    "ofd_solve_tem_adjoint": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int, _vp, _ip,
                               _ip, _dp], _c_int),
    "ofd_assemble_nsvof_jac": ([_vp], _c_int),
    "ofd_vjp_nsvof": ([_vp, _vp, _vp, _vp, _vp], _c_int),
    "ofd_vjp_tem_dual": ([_vp, _vp, _vp], _c_int),
    "ofd_spmv_nsvof_t": ([_vp] + [_vp] * 6, _c_int),
    "ofd_solve_nsvof_t": ([_vp, _c_int, _c_int, _c_dbl, _c_dbl, _c_int, _c_int] + [_vp] * 6 +
                          [_ip, _ip, _dp], _c_int),
    "ofd_fringe_interp": ([_vp, ctypes.c_int32, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ofd_fringe_interp_noalloc": ([_vp, ctypes.c_int32, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "ofd_apply_fringe": ([_vp, _c_int, _vp, _vp], _c_int),
    "ofd_time_spmv": ([_vp, _c_int, _c_int, _dp], _c_int),
    "ofd_time_assemble": ([_vp, _c_int, _c_int, _dp], _c_int),
    "ofd_time_copy": ([_vp, ctypes.c_int64, _c_int, _dp], _c_int),
    "ofd_time_read": ([_vp, ctypes.c_int64, _c_int, _dp], _c_int),
    "ofd_time_quench": ([_vp, _c_int, _c_int, _c_int, _dp], _c_int),
    "ofd_time_adjoint": ([_vp, _c_int, _c_int, _dp], _c_int),
    "ofd_set_weak_faces": ([_vp, _c_int, _vp, _vp], _c_int),
    "ofd_time_faces": ([_vp, _c_int, _c_int, _dp], _c_int),
    "ofd_probe_set": ([_vp, _c_int], _c_int),
    "ofd_probe_read": ([_vp, _i64p, _dp], _c_int),
    "ofd_ctx_bytes": ([_vp, _i64p], _c_int),
}

_LIB = None


class OfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ofd error {code}: {msg}")
        self.code = code


def header_symbols(path=HEADER):
    """Function names declared in include/ofd.h."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ofd_[a-z0-9_]+)\s*\(", txt)))


def load(path=LIB_PATH):
    """Load libofd.so and set every signature; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(
            f"libofd.so not built at {path}: run __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _LIB = lib
    return lib


def check(rc):
    if rc != OFD_OK:
        msg = load().ofd_last_error()
        raise OfdError(rc, msg.decode() if msg else "")
    return rc
