"""ctypes binding of libxg.so (include/xg.h).

This is the stub a Python caller adds to reach the C-ABI; it loads the
in-tree library and fails loudly when it is missing -- there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libxg.so")

XG_OK, XG_EINVAL, XG_ENOMESH, XG_EHIP, XG_EUNSUPPORTED, XG_EBLOWUP, XG_EIO = range(7)
_NAMES = {1: "XG_EINVAL", 2: "XG_ENOMESH", 3: "XG_EHIP", 4: "XG_EUNSUPPORTED",
          5: "XG_EBLOWUP", 6: "XG_EIO"}

# every symbol include/xg.h declares, with (restype, argtypes)
I64, I32, D, VP = C.c_int64, C.c_int32, C.c_double, C.c_void_p
SIG = {
    "xg_ctx_create": (C.c_int, [C.c_int, C.POINTER(VP)]),
    "xg_ctx_destroy": (C.c_int, [VP]),
    "xg_last_error": (C.c_char_p, [VP]),
    "xg_set_stream": (C.c_int, [VP, VP]),
    "xg_set_config": (C.c_int, [VP, VP]),
    "xg_mesh_vol": (C.c_int, [I64, VP, VP, VP, I64, VP, VP]),
    "xg_mesh_domint": (C.c_int, [I64, VP, VP, VP, I64, VP, I64, C.POINTER(I64), VP, VP]),
    "xg_mesh_chocg_domint": (C.c_int, [I64, VP, VP, VP, I64, VP, I64, C.POINTER(I64), VP, VP]),
    "xg_mesh_domsuped": (C.c_int, [I64, VP, I64, VP, VP, C.c_int, VP, VP, VP, VP, VP, VP, VP]),
    "xg_mesh_rcm": (C.c_int, [I64, I64, VP, VP]),
    "xg_mesh_block_order": (C.c_int, [I64, VP, VP, VP, I64, VP]),
    "xg_mesh_symbc": (C.c_int, [I64, VP, VP, VP, I64, VP, VP, I32, VP, I64,
                                C.POINTER(I64), VP, VP, VP]),
    "xg_mesh_neumann": (C.c_int, [I64, VP, VP, VP, I64, VP, VP, VP, VP]),
    "xg_mesh_upload": (C.c_int, [VP, I64, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                 VP, I64, VP, VP]),
    "xg_riemann_grad": (C.c_int, [VP, C.c_int, VP, VP]),
    "xg_riemann_grad_dev": (C.c_int, [VP, C.c_int, VP, VP]),
    "xg_riemann_rhs": (C.c_int, [VP, C.c_int, VP, VP, VP, D, VP, VP]),
    "xg_riemann_rhs_dev": (C.c_int, [VP, C.c_int, VP, VP, VP, D, VP, VP]),
    "xg_set_problem": (C.c_int, [VP, VP]),
    "xg_get_problem": (C.c_int, [VP, VP]),
    "xg_riecg_setup": (C.c_int, [VP, VP, I64, VP, VP]),
    "xg_riecg_set_ncomp": (C.c_int, [VP, C.c_int]),
    "xg_riecg_dt_dev": (C.c_int, [VP, VP, D, C.POINTER(D)]),
    "xg_riecg_stage_dev": (C.c_int, [VP, C.c_int, D, D, VP, VP, VP, VP, VP, VP]),
    "xg_riecg_bc": (C.c_int, [VP, I64, VP, VP, VP, I64, VP, VP, VP, I64, VP, VP]),
    "xg_riecg_apply_bc_dev": (C.c_int, [VP, VP]),
    "xg_riecg_dt_steady_dev": (C.c_int, [VP, VP, D, VP, C.POINTER(D)]),
    "xg_riecg_grad_div_dev": (C.c_int, [VP, VP]),
    "xg_riecg_solve_dev": (C.c_int, [VP, C.c_int, D, D, VP, VP, VP, VP, VP]),
    "xg_diag_err_dev": (C.c_int, [VP, I64, C.c_int, VP, VP, VP, VP]),
    "xg_diag_pre_nc_dev": (C.c_int, [VP, I64, C.c_int, VP, VP, VP, VP, VP, VP]),
    "xg_comm_create": (C.c_int, [C.c_int, C.c_int, VP, VP, C.POINTER(VP)]),
    "xg_comm_sum_dev": (C.c_int, [VP, C.c_int, VP, VP]),
    "xg_comm_shared": (C.c_int, [VP, C.POINTER(I64)]),
    "xg_comm_error": (C.c_char_p, [VP]),
    "xg_comm_destroy": (C.c_int, [VP]),
    "xg_pcg_create": (C.c_int, [C.c_int, C.c_int, C.c_int, VP, VP, VP, VP, VP, C.c_int,
                                C.POINTER(VP)]),
    "xg_pcg_set_stream": (C.c_int, [VP, VP]),
    "xg_pcg_setup": (C.c_int, [VP, VP, VP, C.c_int, C.POINTER(D)]),
    "xg_pcg_solve": (C.c_int, [VP, I64, D, C.POINTER(D), C.POINTER(I64), C.POINTER(C.c_int)]),
    "xg_pcg_solution": (C.c_int, [VP, C.c_int, VP]),
    "xg_pcg_error": (C.c_char_p, [VP]),
    "xg_pcg_destroy": (C.c_int, [VP]),
    "xg_mesh_device_bytes": (C.c_int, [VP, C.POINTER(I64)]),
    "xg_kozak_upload": (C.c_int, [VP, I64, VP, VP, VP, I64, VP]),
    "xg_kozak_rhs": (C.c_int, [VP, C.c_int, D, D, VP, VP, VP, VP]),
    "xg_kozak_rhs_dev": (C.c_int, [VP, C.c_int, D, D, VP, VP, VP, VP]),
    "xg_kozcg_setup": (C.c_int, [VP, VP, I64, VP, VP, I64, VP, VP, VP, VP]),
    "xg_chocg_setup": (C.c_int, [VP, VP, VP, I64, VP, VP, I64, VP, I64, VP, VP, VP,
                                 I64, VP, VP]),
    "xg_chocg_initial_dev": (C.c_int, [VP, VP, VP, VP]),
    "xg_chocg_set_ncomp": (C.c_int, [VP, C.c_int]),
    "xg_chocg_step2_dev": (C.c_int, [VP, D, C.c_int, VP, VP, VP, VP, VP]),
    "xg_chocg_poisson_dev": (C.c_int, [VP, VP, VP, VP, VP, VP]),
    "xg_chocg_dt_dev": (C.c_int, [VP, VP, VP]),
    "xg_chocg_step_dev": (C.c_int, [VP, D, VP, VP, VP, VP, VP]),
    "xg_chocg_momentum_setup": (C.c_int, [VP, I64, VP, D, C.c_int, I64, D]),
    "xg_chocg_momentum_its": (C.c_int, [VP, VP]),
    "xg_cg_init_dev": (C.c_int, [VP, VP, I64, VP, VP, VP, C.c_int, C.c_int]),
    "xg_diag_rho_dev": (C.c_int, [VP, I64, C.c_int, VP, VP, VP, VP]),
    "xg_diag_rho_dev_noalloc": (C.c_int, [VP, I64, C.c_int, VP, VP, VP, VP]),
    "xg_amg_info": (C.c_int, [VP, VP, VP, VP]),
    "xg_diag_pre_dev": (C.c_int, [VP, I64, VP, VP, VP, VP, VP, VP]),
    "xg_kozcg_dt_dev": (C.c_int, [VP, VP, D, VP]),
    "xg_kozcg_step_dev": (C.c_int, [VP, D, D, VP, VP]),
    "xg_kozcg_set_ncomp": (C.c_int, [VP, C.c_int]),
    "xg_kozcg_dt_steady_dev": (C.c_int, [VP, VP, D, VP, VP]),
    "xg_kozcg_step2_dev": (C.c_int, [VP, D, D, VP, VP, C.c_int, VP, VP]),
    "xg_exo_error": (C.c_char_p, [VP]),
    "xg_exo_close": (None, [VP]),
    "xg_exo_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(VP)]),
    "xg_exo_sizes": (C.c_int, [VP, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                               C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(I64)]),
    "xg_exo_coords": (C.c_int, [VP, VP, VP, VP]),
    "xg_exo_tets": (C.c_int, [VP, VP]),
    "xg_exo_tris": (C.c_int, [VP, VP]),
    "xg_exo_sideset": (C.c_int, [VP, C.c_int, C.POINTER(C.c_int), C.POINTER(I64)]),
    "xg_exo_sideset_faces": (C.c_int, [VP, C.c_int, VP]),
    "xg_exo_node_var_name": (C.c_int, [VP, C.c_int, C.c_char_p, C.c_int]),
    "xg_exo_times": (C.c_int, [VP, VP]),
    "xg_exo_read_node_var": (C.c_int, [VP, I64, C.c_int, VP]),
    "xg_exo_create": (C.c_int, [C.c_char_p, I64, VP, VP, VP, I64, VP, I64, VP, C.c_int, VP,
                                VP, VP, C.c_int, VP, C.POINTER(VP)]),
    "xg_exo_write_time": (C.c_int, [VP, I64, D]),
    "xg_exo_write_node_var": (C.c_int, [VP, I64, C.c_int, VP]),
    "xg_chocg_upload": (C.c_int, [VP, I64, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                  I64, VP]),
    "xg_chorin_div": (C.c_int, [VP, C.c_int, D, VP, VP, C.c_int, VP, C.c_int, VP,
                                C.c_int]),
    "xg_chorin_grad": (C.c_int, [VP, C.c_int, VP, VP, C.c_int]),
    "xg_chorin_vgrad": (C.c_int, [VP, C.c_int, C.c_int, VP, VP]),
    "xg_chorin_flux": (C.c_int, [VP, C.c_int, D, VP, C.c_int, VP, C.c_int, VP]),
    "xg_chorin_rhs": (C.c_int, [VP, C.c_int, VP, C.c_int, VP, VP, VP, VP]),
    "xg_csr_pattern": (C.c_int, [I64, I64, VP, C.c_int, I64, C.POINTER(I64), VP, VP]),
    "xg_poisson_matrix": (C.c_int, [I64, VP, VP, VP, I64, VP, VP, VP, VP]),
    "xg_cg_matrix": (C.c_int, [VP, C.c_int, I64, VP, VP, VP, C.c_int]),
    "xg_cg_init": (C.c_int, [VP, VP, VP, VP, I64, VP, VP, VP, C.c_int, C.c_int]),
    "xg_cg_setup": (C.c_int, [VP, C.POINTER(D)]),
    "xg_cg_solve": (C.c_int, [VP, I64, D, C.POINTER(D), C.POINTER(I64),
                              C.POINTER(C.c_int)]),
    "xg_cg_solution": (C.c_int, [VP, VP]),
    "xg_cg_device_x": (C.c_int, [VP, C.POINTER(VP)]),
    "xg_cg_solve_mixed": (C.c_int, [VP, I64, D, D, I64, C.POINTER(D), C.POINTER(I64),
                                    C.POINTER(I64), C.POINTER(C.c_int)]),
    "xg_cg_solve_reliable": (C.c_int, [VP, I64, D, D, C.POINTER(D), C.POINTER(I64),
                                       C.POINTER(I64), C.POINTER(C.c_int)]),
    "xg_cg_solve_flexible": (C.c_int, [VP, C.c_int, I64, I64, D, C.POINTER(D),
                                       C.POINTER(I64), C.POINTER(I64), C.POINTER(C.c_int)]),
    "xg_gnn_create": (C.c_int, [VP, I64, I64]),
    "xg_gnn_set_params": (C.c_int, [VP, VP, I64]),
    "xg_gnn_get_params": (C.c_int, [VP, VP, VP, I64]),
    "xg_gnn_adam": (C.c_int, [VP, D, D, D, D]),
    "xg_gnn_prepare": (C.c_int, [VP]),
    "xg_gnn_apply_dev": (C.c_int, [VP, VP, VP]),
    "xg_gnn_loss_grad_dev": (C.c_int, [VP, VP, VP, VP, C.POINTER(D)]),
    "xg_gnn_train_system": (C.c_int, [VP, I64, I64, C.c_int, D, C.POINTER(D),
                                      C.POINTER(I64)]),
    "xg_cg_destroy": (C.c_int, [VP]),
}


class XgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_NAMES.get(code, code), msg))
        self.code = code


class Config(C.Structure):
    _fields_ = [("gamma", C.c_double), ("flux", C.c_int), ("stab2", C.c_int),
                ("stab2coef", C.c_double)]


class Problem(C.Structure):
    _fields_ = [("kind", C.c_int), ("alpha", C.c_double), ("kappa", C.c_double),
                ("p0", C.c_double)]


PROBLEMS = {"none": 0, "vortical_flow": 1, "taylor_green": 2, "slot_cyl": 3}


class ChoConfig(C.Structure):
    _fields_ = [("flux", C.c_int), ("stab", C.c_int), ("stab2", C.c_int),
                ("stab2coef", C.c_double), ("mu", C.c_double), ("diff", C.c_double)]


class ChoStep(C.Structure):
    _fields_ = [("flux", ChoConfig), ("cfl", C.c_double), ("rk", C.c_int),
                ("p_pc", C.c_int), ("p_maxit", C.c_int64), ("p_tol", C.c_double)]


class FctConfig(C.Structure):
    _fields_ = [("fct", C.c_int), ("fctdif", C.c_double), ("fctclip", C.c_int),
                ("nsys", C.c_int), ("sys", C.c_int * 5)]


_LIB = None


def lib():
    """Load libxg.so (built by __graft_entry__.build / make in csrc)."""
    global _LIB
    if _LIB is None:
        # PyTorch bundles its own HIP runtime (soname libamdhip64.so.7).  Load
        # it first so that libxg.so binds to the same runtime: two HIP/HSA
        # runtimes in one process do not share the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIBPATH):
            raise ImportError("xyst_amd: %s is not built (run __graft_entry__.build())"
                              % LIBPATH)
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ptr(a):
    """Address of a C-contiguous numpy array or a torch tensor's storage."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
        return a.ctypes.data
    assert a.is_contiguous(), "tensor must be contiguous"
    return a.data_ptr()


def check(rc, ctx=None, what=""):
    if rc != XG_OK:
        msg = lib().xg_last_error(ctx).decode() if ctx else what
        raise XgError(rc, msg or what)
