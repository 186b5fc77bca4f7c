"""ctypes binding of ``libbecgpp_hip.so`` (the C-ABI declared in include/becgpp_hip.h).

The library is the only compute path: when it is missing or no GPU is
visible, every operator raises instead of falling back to another backend.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbecgpp_hip.so")

KERNEL_NONE, KERNEL_NEWTON, KERNEL_LOG, KERNEL_TABLE = 0, 1, 2, 3
WHICH_K, WHICH_W = 0, 1
CG_BETA = {"pr_precond": 0, "pr": 1, "none": 2}
STOP_REASONS = ("maxit", "converged", "energy_converged", "line_search_stalled")

c_double_p = ctypes.POINTER(ctypes.c_double)


class Phys(ctypes.Structure):
    _fields_ = [("beta2", ctypes.c_double), ("beta3", ctypes.c_double),
                ("Omega", ctypes.c_double), ("G_C", ctypes.c_double)]


class SolverCfg(ctypes.Structure):
    _fields_ = [("maxit", ctypes.c_int), ("res_tol", ctypes.c_double),
                ("energy_tol", ctypes.c_double), ("conv_window", ctypes.c_int),
                ("step", ctypes.c_double), ("step_max", ctypes.c_double),
                ("linesearch_max", ctypes.c_int), ("linesearch_stall", ctypes.c_int),
                ("precond_shift_min", ctypes.c_double), ("cg_beta", ctypes.c_int),
                ("cg_restart", ctypes.c_int), ("verbose", ctypes.c_int), ("check", ctypes.c_int),
                ("host_loop", ctypes.c_int), ("warm_res_tol", ctypes.c_double), ("line_poly", ctypes.c_int)]


class SeedSpec(ctypes.Structure):
    _fields_ = [("envelope", ctypes.c_int), ("amp_noise", ctypes.c_double), ("phase_noise", ctypes.c_double),
                ("winding", ctypes.c_int), ("vortex_sign", ctypes.c_double), ("core", ctypes.c_double),
                ("rng", ctypes.c_ulonglong)]


class SolveOut(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("converged", ctypes.c_int),
                ("stop_reason", ctypes.c_int), ("n_trace", ctypes.c_int),
                ("E", ctypes.c_double), ("resid_rel", ctypes.c_double),
                ("best_resid", ctypes.c_double), ("sigma", ctypes.c_double),
                ("walltime", ctypes.c_double), ("n_energy", ctypes.c_longlong),
                ("warm_iters", ctypes.c_int)]


# name -> (restype, argtypes); every entry point returns an int status
SIGNATURES = {
    "bec_last_error": (ctypes.c_char_p, []),
    "bec_plan_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "bec_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "bec_plan_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "bec_plan_build_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "bec_plan_set_kernel_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "bec_plan_clear_kernels": (ctypes.c_int, [ctypes.c_void_p]),
    "bec_norm2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_double_p]),
    "bec_conv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double,
                                ctypes.c_void_p]),
    "bec_energy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Phys), c_double_p,
                                  ctypes.c_void_p]),
    "bec_apply_H": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Phys),
                                   ctypes.c_void_p, ctypes.c_void_p]),
    "bec_Lz": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bec_residual": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Phys), c_double_p]),
    "bec_precondition": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(Phys), ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_void_p, c_double_p]),
    "bec_moments": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_double_p]),
    "bec_lll_weight": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, c_double_p,
                                      ctypes.POINTER(ctypes.c_int)]),
    "bec_observe_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Phys), ctypes.c_void_p,
                                         c_double_p]),
    "bec_lll_weight_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, c_double_p, ctypes.c_int,
                                            c_double_p]),
    "bec_vortex_count_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, c_double_p,
                                              ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    "bec_plan_max_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "bec_resample": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                    ctypes.c_void_p]),
    "bec_resample_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]),
    "bec_vortex_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_int)]),
    "bec_tf_mu": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, c_double_p]),
    "bec_tf_nonlocal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_void_p, c_double_p]),
    "bec_seed": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SeedSpec), ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_int, ctypes.c_void_p]),
    "bec_noise": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_longlong,
                                 c_double_p]),
    "bec_ground_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Phys), ctypes.POINTER(SolverCfg),
                                        ctypes.c_void_p, ctypes.POINTER(SolveOut), c_double_p,
                                        ctypes.c_int]),
    "bec_ground_state_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Phys),
                                              ctypes.POINTER(SolverCfg), ctypes.c_void_p,
                                              ctypes.POINTER(SolveOut)]),
    "bec_ground_state_queue": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Phys),
                                              ctypes.POINTER(SolverCfg), ctypes.c_void_p, ctypes.POINTER(SolveOut)]),
    "bec_ground_state_uneven": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Phys),
                                               ctypes.POINTER(ctypes.POINTER(SolverCfg)),
                                               ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(SolveOut)]),
    "bec_plan_reserve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "bec_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_double_p, c_double_p, ctypes.c_char_p,
                                         ctypes.c_int]),
    "bec_prof_watch": (ctypes.c_int, [ctypes.c_int]),
    "bec_prof_read": (ctypes.c_int, [c_double_p, ctypes.POINTER(ctypes.c_longlong), c_double_p]),
    "bec_prof_sites": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "bec_prof_read_site": (ctypes.c_int, [ctypes.c_int, c_double_p, ctypes.POINTER(ctypes.c_longlong), c_double_p]),
}

_LIB = None


def load():
    """Load the library (once) and declare its signatures; raise if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m becgpp_amd._build` "
                           "(there is no fallback compute path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(status):
    if status == 0:
        return
    msg = load().bec_last_error().decode(errors="replace")
    if status == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def phys(beta2, beta3, Omega, G_C):
    return Phys(float(beta2), float(beta3), float(Omega), float(G_C))
