"""Loader of libfcp.so (the HIP path).  There is no fallback: if the library
is missing or does not load, every call raises."""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# FCP_LIB selects another build of the same library (e.g. an instrumented one).
LIB_PATH = os.environ.get("FCP_LIB") or os.path.join(PKG_DIR, "lib", "libfcp.so")

_i = C.c_int
_d = C.c_double
_p = C.c_void_p


class SolveOpts(C.Structure):
    _fields_ = [("precision", _i), ("criterion", _i), ("inner_tol", _d), ("max_outer", _i),
                ("check_every", _i), ("iterations", _i), ("outer_iterations", _i),
                ("res0", _d), ("resl", _d), ("resor", _d), ("rel_res_l2", _d), ("time_ms", _d)]


# name -> (restype, argtypes); mirrors include/fcp.h
SIGNATURES = {
    "fcp_init": (_i, [_i, C.POINTER(_p)]),
    "fcp_finalize": (None, [_p]),
    "fcp_last_error": (C.c_char_p, [_p]),
    "fcp_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "fcp_mesh_upload": (_i, [_p, _i, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p]),
    "fcp_geometry": (_i, [_p] + [_p] * 14),
    "fcp_read_polymesh": (_i, [_p, C.c_char_p, _i, _i, _p]),
    "fcp_mesh_periodic": (_i, [_p, _i, _p]),
    "fcp_lis_create_solver": (_i, [_p, C.c_char_p, _i]),
    "fcp_lis_spsolve": (_i, [_p, _p, _p, C.c_char_p, _i, _p, _p, _p]),
    "fcp_lis_destroy_solver": (_i, [_p]),
    "fcp_mesh_patches": (_i, [_p, _p, _p, _p]),
    "fcp_csr_build": (_i, [_p, C.POINTER(_i), _p, _p, _p, _p, _p]),
    "fcp_levels": (_i, [_p, _i, _p, C.POINTER(_i)]),
    "fcp_laplacian": (_i, [_p, _p, _p, _p, _p]),
    "fcp_laplacian_face": (_i, [_p, _p, _p, _p, _p, _p]),
    "fcp_nonorth_correction": (_i, [_p, _p, _p, _p]),
    "fcp_grad_gauss": (_i, [_p, _p, _p]),
    "fcp_convdiff": (_i, [_p, _p, _p, _p, _p, _d, _i, _p, _d, _p, _p, _p]),
    "fcp_convdiff_cell": (_i, [_p, _p, _p, _p, _p, _d, _i, _p, _d, _p, _p, _p]),
    "fcp_momentum_assemble": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                   _d, _d, _i, _i, _d, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "fcp_pressure_assemble": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _d, _p, _p]),
    "fcp_pressure_correct": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _d, _i, _p, _p,
                                  _p]),
    "fcp_ddt": (_i, [_p, _i, _p, _d, _p, _p, _p, _p]),
    "fcp_transport": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _d, _d, _d, _d, _d, _d, _i, _i, _i,
                           C.c_char_p, _i, _i, _d, _d, _i, _p, _p, _p, _p, C.POINTER(_d),
                           C.POINTER(_d), C.POINTER(_d), C.POINTER(SolveOpts)]),
    "fcp_update_boundary": (_i, [_p, _p]),
    "fcp_matrix_set": (_i, [_p, _p]),
    "fcp_spmv": (_i, [_p, _p, _p]),
    "fcp_precond_apply": (_i, [_p, _i, _p, _p, _p]),
    "fcp_spmv_transpose": (_i, [_p, _p, _p]),
    "fcp_ilu_factor": (_i, [_p, _p]),
    "fcp_sweep_tiles": (_i, [_p, _p]),
    "fcp_autotune": (_i, [_p, C.c_char_p, _i, C.POINTER(_i), C.POINTER(_i)]),
    "fcp_csrsolve": (_i, [_p, C.c_char_p, _i, _p, _p, C.POINTER(_d), _i, _d, _d, C.c_char_p, _i,
                          _i, C.POINTER(SolveOpts)]),
    "fcp_rcm": (_i, [_p, _i, _p]),
    "fcp_matrix_transpose": (_i, [_p]),
    "fcp_convdiff_backward": (_i, [_p, _p, _p, _p, _p, _d, _i, _p, _p, _p, _p, _p, _p]),
    "fcp_laplacian_face_backward": (_i, [_p] + [_p] * 11),
    "fcp_grad_gauss_backward": (_i, [_p, _p, _p]),
    "fcp_create_lsq_grad_matrix": (_i, [_p, _i, _i, _p]),
    "fcp_grad": (_i, [_p, _i, _p, _p]),
    "fcp_csrsolve_multi": (_i, [_p, C.c_char_p, _i, _i, _p, _i, _p, _p, _i, _d, _d, _i,
                                C.POINTER(SolveOpts)]),
    "fcp_csrsolve_multi_diag": (_i, [_p, C.c_char_p, _i, _i, _p, _i, _p, _p, _p, _i, _d, _d, _i,
                                     C.POINTER(SolveOpts)]),
    "fcp_iccg": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, C.POINTER(_d), _i, _d, _d, _i]),
    "fcp_dpcg": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, C.POINTER(_d), _i, _d, _d, _i]),
    "fcp_bicgstab": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, C.POINTER(_d), _i, _d, _d, _i]),
    "fcp_gauss_seidel": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, C.POINTER(_d), _i, _d, _d, _i]),
    "fcp_pmgmres": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, C.POINTER(_d), _i, _i, _d, _d, _i]),
    "fcp_dpcg_noalloc": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, C.POINTER(_d), _i, _d, _d, _p]),
    "fcp_profile": (_i, [_p, _i]),
    "fcp_profile_report": (_i, [_p, C.c_char_p, _i]),
    "fcp_stream": (_p, [_p]),
    "fcp_tune_set": (_i, [_p, _i, _i]),
}

_lib = None


def load():
    """Load libfcp.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libfcp.so not found at {LIB_PATH}; run __graft_entry__.build()")
        # PyTorch ships its own HIP and HSA runtime libraries.  Loaded first,
        # they also serve libfcp (same sonames), so both share one runtime and
        # exchange device pointers; loaded after libfcp's, torch finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
