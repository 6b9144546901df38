"""ctypes binding of libh4hip.so (include/h4hip.h).

The library is loaded from the package directory only; if it is missing
the import fails loudly -- there is no CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libh4hip.so")

H4H_OK = 0
H4H_EINVAL = -1
H4H_EHIP = -2
H4H_ENOMEM = -3
H4H_ESTATE = -4
H4H_ENODEV = -5

H4H_FP64 = 0
H4H_FP64_EXACT = 1
H4H_MIXED = 2

PRECISIONS = {"fp64": H4H_FP64, "exact": H4H_FP64_EXACT, "mixed": H4H_MIXED}

# Every symbol include/h4hip.h declares (checked by the CPU test suite).
EXPORTS = (
    "h4h_abi_version", "h4h_last_error", "h4h_create", "h4h_destroy",
    "h4h_set_precision", "h4h_set_launch", "h4h_upload", "h4h_download",
    "h4h_download_predictor", "h4h_upload_predictor", "h4h_init_acc_jrk",
    "h4h_init_dt", "h4h_find_particles_to_move", "h4h_set_active",
    "h4h_get_active", "h4h_save_old_acc_jrk", "h4h_predicted_pos_vel",
    "h4h_update_acc_jrk", "h4h_correction_pos_vel",
    "h4h_next_integration_time", "h4h_step", "h4h_run", "h4h_get_energy",
    "h4h_set_timing", "h4h_get_stats", "h4h_synchronize", "h4h_create_ensemble",
    "h4h_ensemble_sizes", "h4h_ensemble_init", "h4h_ensemble_run", "h4h_ensemble_energy",
    "h4h_set_tuning", "h4h_autotune", "h4h_set_graphs", "h4h_set_small_path",
    "h4h_tune_launch",
    "h4h_set_launch_choice", "h4h_update_acc_jrk_static",
)

# Limits of h4h_update_acc_jrk_static (include/h4hip.h).
H4H_STATIC_MAX_N = 65536
H4H_STATIC_SCRATCH_BYTES = 58130688


class H4HStats(ctypes.Structure):
    _fields_ = [
        ("block_steps", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("nsteps", ctypes.c_uint64),
        ("update_launches", ctypes.c_uint64),
        ("update_ms", ctypes.c_double),
        ("step_ms", ctypes.c_double),
        ("small_steps", ctypes.c_uint64),
    ]


class H4HRunInfo(ctypes.Structure):
    _fields_ = [
        ("itime", ctypes.c_double),
        ("atime", ctypes.c_double),
        ("steps", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("nsteps", ctypes.c_uint64),
        ("last_nact", ctypes.c_uint32),
        ("stopped", ctypes.c_int),
    ]


class H4HClusterInfo(ctypes.Structure):
    _fields_ = [
        ("itime", ctypes.c_double),
        ("atime", ctypes.c_double),
        ("steps", ctypes.c_uint64),
        ("interactions", ctypes.c_uint64),
        ("nsteps", ctypes.c_uint64),
        ("stopped", ctypes.c_int),
    ]


class H4HLaunchChoice(ctypes.Structure):
    _fields_ = [
        ("workgroups", ctypes.c_uint32),
        ("max_jsplit", ctypes.c_uint32),
        ("iw_fill16", ctypes.c_uint32),
        ("ms_per_step", ctypes.c_float),
    ]


H4H_STOP_OUTPUT = 1
H4H_STOP_END = 2
H4H_STOP_MAXSTEPS = 3


class H4HError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


_lib = None


def load():
    """Loads libh4hip.so and declares its prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32, i32, dbl, flt = ctypes.c_uint32, ctypes.c_int, ctypes.c_double, ctypes.c_float
    pd = ctypes.POINTER(ctypes.c_double)
    pu = ctypes.POINTER(ctypes.c_uint32)
    pf = ctypes.POINTER(ctypes.c_float)
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "h4h_abi_version": ([], i32),
        "h4h_last_error": ([], ctypes.c_char_p),
        "h4h_create": ([pp, u32, dbl, flt, dbl, dbl, i32], i32),
        "h4h_destroy": ([P], i32),
        "h4h_set_precision": ([P, i32], i32),
        "h4h_set_launch": ([P, u32, u32], i32),
        "h4h_upload": ([P, pd, pd, pd, pd, pd], i32),
        "h4h_download": ([P, pd, pd, pd, pd, pd, pd, pd, pd], i32),
        "h4h_download_predictor": ([P, pd, pf], i32),
        "h4h_upload_predictor": ([P, pd, pf], i32),
        "h4h_init_acc_jrk": ([P], i32),
        "h4h_init_dt": ([P, pd, flt, dbl], i32),
        "h4h_find_particles_to_move": ([P, dbl, pu], i32),
        "h4h_set_active": ([P, pu, u32], i32),
        "h4h_get_active": ([P, pu, pu], i32),
        "h4h_save_old_acc_jrk": ([P], i32),
        "h4h_predicted_pos_vel": ([P, dbl], i32),
        "h4h_update_acc_jrk": ([P, u32], i32),
        "h4h_correction_pos_vel": ([P, dbl, u32], i32),
        "h4h_next_integration_time": ([P, pd], i32),
        "h4h_step": ([P, dbl, pu, pd], i32),
        "h4h_run": ([P, dbl, dbl, ctypes.c_uint64, ctypes.POINTER(H4HRunInfo), pd, pu], i32),
        "h4h_get_energy": ([P, pd, pd], i32),
        "h4h_set_timing": ([P, i32], i32),
        "h4h_get_stats": ([P, ctypes.POINTER(H4HStats), i32], i32),
        "h4h_synchronize": ([P], i32),
        "h4h_create_ensemble": ([pp, u32, pu, dbl, flt, dbl, dbl, i32], i32),
        "h4h_ensemble_sizes": ([P, pu, pu], i32),
        "h4h_ensemble_init": ([P, flt], i32),
        "h4h_ensemble_run": ([P, dbl, dbl, ctypes.c_uint64, ctypes.POINTER(H4HClusterInfo)], i32),
        "h4h_ensemble_energy": ([P, pd, pd], i32),
        "h4h_set_tuning": ([P, pu, pu], i32),
        "h4h_autotune": ([P, u32, pu, pu], i32),
        "h4h_set_graphs": ([P, i32], i32),
        "h4h_set_small_path": ([P, u32], i32),
        "h4h_tune_launch": ([P, u32, ctypes.POINTER(H4HLaunchChoice)], i32),
        "h4h_set_launch_choice": ([P, ctypes.POINTER(H4HLaunchChoice)], i32),
        # device pointers as integers (e.g. a torch tensor's data_ptr())
        "h4h_update_acc_jrk_static": ([P, P, u32, P, u32, dbl, i32, P, P], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(func, rc):
    if rc != H4H_OK:
        msg = load().h4h_last_error().decode(errors="replace")
        raise H4HError(func, rc, msg)
    return rc
