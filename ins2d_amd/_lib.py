"""ctypes binding of libins2d_hip.so (include/ins2d.h).

The library must have been built in-tree (``python -m ins2d_amd.build``);
loading fails loudly otherwise -- there is no CPU fallback."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libins2d_hip.so")

c_int, c_double, c_void_p, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
P_double = ctypes.POINTER(ctypes.c_double)
P_int = ctypes.POINTER(ctypes.c_int)
P_float = ctypes.POINTER(ctypes.c_float)


class CaseDesc(ctypes.Structure):
    _fields_ = [("Nx", c_int), ("Ny", c_int), ("x", P_double), ("y", P_double),
                ("bc_u", c_int * 4), ("bc_v", c_int * 4), ("Re", c_double), ("order4", c_int),
                ("alfa", c_double), ("beta", c_double), ("rk_stages", c_int), ("rk_A", P_double),
                ("rk_b", P_double), ("rk_c", P_double), ("poisson", c_int), ("fft_symbol", c_int),
                ("cg_tol", c_double), ("cg_maxit", c_int), ("p_add_solve", c_int), ("kernels", c_int),
                ("cg_precond", c_int)]


# every entry point declared in include/ins2d.h: name -> (restype, argtypes)
SIGNATURES = {
    "ins2d_create": (c_int, [ctypes.POINTER(CaseDesc), ctypes.POINTER(c_void_p)]),
    "ins2d_destroy": (None, [c_void_p]),
    "ins2d_last_error": (c_char_p, [c_void_p]),
    "ins2d_grid_info": (c_int, [c_void_p, P_int, P_double, P_double]),
    "ins2d_set_bc": (c_int, [c_void_p, c_double, ctypes.POINTER(P_double)]),
    "ins2d_set_state": (c_int, [c_void_p, P_double, P_double]),
    "ins2d_get_state": (c_int, [c_void_p, P_double, P_double]),
    "ins2d_rhs": (c_int, [c_void_p, P_double, P_double, c_double, P_double, P_double]),
    "ins2d_poisson": (c_int, [c_void_p, P_double, P_double, P_int, P_double]),
    "ins2d_input_check": (c_int, [c_void_p, c_double, c_int]),
    "ins2d_step": (c_int, [c_void_p, c_double, c_double, c_int]),
    "ins2d_set_bc_rate": (c_int, [c_void_p, ctypes.POINTER(P_double)]),
    "ins2d_step_unsteady": (c_int, [c_void_p, c_double, c_double, c_int, P_double, P_double]),
    "ins2d_diagnostics": (c_int, [c_void_p, c_double, P_double]),
    "ins2d_solver_stats": (c_int, [c_void_p, P_int, P_double, ctypes.POINTER(ctypes.c_longlong)]),
    "ins2d_profile": (c_int, [c_void_p, c_int]),
    "ins2d_profile_read": (c_int, [c_void_p, c_int, P_double, ctypes.POINTER(ctypes.c_longlong)]),
    "ins2d_sync": (c_int, [c_void_p]),
    "ins2d_batch_create": (c_int, [ctypes.POINTER(CaseDesc), c_int, ctypes.POINTER(c_void_p)]),
    "ins2d_batch_destroy": (None, [c_void_p]),
    "ins2d_batch_last_error": (c_char_p, [c_void_p]),
    "ins2d_batch_info": (c_int, [c_void_p, c_int, P_int, P_double, P_double]),
    "ins2d_batch_set_bc": (c_int, [c_void_p, c_int, c_double, ctypes.POINTER(P_double)]),
    "ins2d_batch_set_state": (c_int, [c_void_p, c_int, P_double, P_double]),
    "ins2d_batch_get_state": (c_int, [c_void_p, c_int, P_double, P_double]),
    "ins2d_batch_input_check": (c_int, [c_void_p, c_int]),
    "ins2d_batch_step": (c_int, [c_void_p, P_double, c_int]),
    "ins2d_batch_stats": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_longlong), P_int, ctypes.POINTER(ctypes.c_longlong)]),
    "ins2d_batch_autotune": (c_int, [c_void_p, P_int, P_int]),
    "ins2d_batch_set_tile": (c_int, [c_void_p, c_int, c_int]),
    "ins2d_nn_create": (c_int, [c_int, c_int, ctypes.POINTER(c_void_p)]),
    "ins2d_nn_destroy": (None, [c_void_p]),
    "ins2d_nn_last_error": (c_char_p, [c_void_p]),
    "ins2d_nn_num_params": (ctypes.c_long, [c_void_p]),
    "ins2d_nn_set_params": (c_int, [c_void_p, P_float, c_double]),
    "ins2d_nn_get_params": (c_int, [c_void_p, P_float]),
    "ins2d_nn_init": (c_int, [c_void_p, ctypes.c_ulonglong, c_double]),
    "ins2d_nn_forward": (c_int, [c_void_p, P_double, P_double, P_double]),
    "ins2d_nn_forward_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ins2d_nn_forward_unfused": (c_int, [c_void_p, P_double, P_double, P_double]),
    "ins2d_nn_forward_unfused_dev": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ins2d_nn_train_step": (c_int, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                    ctypes.POINTER(c_void_p), c_int, ctypes.c_float, c_int, P_double, c_void_p]),
    "ins2d_nn_get_grad": (c_int, [c_void_p, P_float]),
    "ins2d_nn_data_reserve": (c_int, [c_void_p, c_int]),
    "ins2d_nn_data_set": (c_int, [c_void_p, c_int, P_double, P_double, P_double]),
    "ins2d_nn_train": (c_int, [c_void_p, P_int, c_int, ctypes.c_float, c_int, P_double]),
    "ins2d_nn_layer_check": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, P_float, P_float, P_float,
                                     P_float, P_float, P_float, P_float, P_float, P_float]),
    "ins2d_nn_mfma_probe": (c_int, [P_float, P_float, P_float]),
    "ins2d_nn_shape": (c_int, [c_void_p, P_int, P_int]),
    "ins2d_set_poisson_accel": (c_int, [c_void_p, c_void_p, c_int, c_int]),
    "ins2d_poisson_accel_stats": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_longlong), P_int]),
    "ins2d_poisson_coarse": (c_int, [c_void_p, P_double, P_double]),
    "ins2d_pressure_rhs_dev": (c_int, [c_void_p, c_double, c_void_p]),
    "ins2d_poisson_dev": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ins2d_poisson_coarse_dev": (c_int, [c_void_p, c_void_p, c_void_p]),
    "ins2d_nn_data_ptr": (c_void_p, [c_void_p, c_int, c_int]),
    "ins2d_nn_data_get": (c_int, [c_void_p, c_int, P_double, P_double, P_double]),
    "ins2d_nn_data_combine": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double]),
    "ins2d_pressure_rhs": (c_int, [c_void_p, c_double, P_double]),
    "ins2d_residual": (c_int, [c_void_p, P_double, P_double, P_double]),
    "ins2d_fdm_mfma_probe": (c_int, [P_double, P_double, P_double]),
    "ins2d_poisson_fft_noalloc": (c_int, [c_int, c_int, c_double, c_double, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p]),
}

_LIB = None


def load(path=LIB_PATH):
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"libins2d_hip is not built ({path} missing): run `python -m ins2d_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
