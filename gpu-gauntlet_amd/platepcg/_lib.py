"""ctypes binding of the platepcg C ABI (include/platepcg.h).

The library is the in-tree build ``gpu-gauntlet_amd/lib/libplatepcg.so``.
There is no fallback: if it is missing or fails to load, ``load()`` raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libplatepcg.so")

PP_OK = 0
PP_ERR_INVALID = -1
PP_ERR_HIP = -2
PP_ERR_UNSUPPORTED = -3
PP_ERR_NOMEM = -4
PP_ERR_TIMEOUT = -5
PP_ERR_NODEVICE = -6

PP_FIX_W, PP_FIX_THX, PP_FIX_THY = 1, 2, 4

PRECISION_FP64, PRECISION_MIXED = 0, 1
PRECOND_BLOCK_JACOBI, PRECOND_LEARNED, PRECOND_CHEBYSHEV, PRECOND_MULTIGRID = 0, 1, 2, 3
PATH_AUTO, PATH_STRUCTURED, PATH_COLOURED = 0, 1, 2
MAX_SWEEPS = 32
FLAG_NO_TRUE_RESIDUAL = 1
FLAG_SMALL_TILES = 2
FLAG_CLASSIC_PCG = 4
FLAG_RELIABLE_UPDATE = 8
FLAG_MG_UNFUSED = 16
FLAG_COLOURED_HOST_LOOP = 32


class Mesh(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("elements", C.c_void_p), ("node_count", C.c_uint32),
                ("element_count", C.c_uint32), ("nodes_per_element", C.c_uint32),
                ("grid_nx", C.c_uint32), ("grid_ny", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("t", C.c_double),
                ("element_t", C.c_void_p)]


class Colouring(C.Structure):
    _fields_ = [("elements_by_colour", C.c_void_p), ("colour_offsets", C.c_void_p),
                ("colour_count", C.c_uint32)]


class SolveOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int32),
                ("precision", C.c_int32), ("precond", C.c_int32), ("path", C.c_int32),
                ("inner_tolerance", C.c_double), ("max_refinements", C.c_int32),
                ("flags", C.c_int32), ("poly_degree", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("breakdown", C.c_int32),
                ("sweeps", C.c_int32), ("final_residual", C.c_double),
                ("true_rel_residual", C.c_double), ("device_ms", C.c_double),
                ("sweep_iterations", C.c_int32 * MAX_SWEEPS),
                ("sweep_residual", C.c_double * MAX_SWEEPS), ("phase_us", C.c_double * 6)]


TRAIN_CURVE = 64
TRAIN_SYNTHETIC, TRAIN_SOLVER, TRAIN_LEARNED = 0, 1, 2


class CnnConfig(C.Structure):
    _fields_ = [("channels", C.c_int32 * 3), ("bottleneck_dilation", C.c_int32),
                ("rotation_scale", C.c_double), ("smoother_weight", C.c_double),
                ("relu", C.c_int32), ("fixed_step", C.c_int32),
                ("fcg_window", C.c_int32), ("multigrid", C.c_int32), ("mg_sweeps", C.c_int32),
                ("mg_coarse_sweeps", C.c_int32), ("mg_jacobi", C.c_int32),
                ("unfused_forward", C.c_int32)]


class TrainOpts(C.Structure):
    _fields_ = [("pairs", C.c_int32), ("epochs", C.c_int32), ("seed0", C.c_uint64),
                ("source", C.c_int32), ("snapshots", C.c_int32),
                ("lr", C.c_double), ("lr_final", C.c_double), ("max_grad_norm", C.c_double)]


class TrainStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("curve_points", C.c_int32), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("seconds", C.c_double),
                ("loss_curve", C.c_double * TRAIN_CURVE)]


class BatchSchedule(C.Structure):
    _fields_ = [("workgroups_per_cu", C.c_int32), ("tile_shape", C.c_int32),
                ("balance_iterations", C.c_int32), ("resident_plates", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("plates", C.c_int32), ("converged", C.c_int32), ("workgroups", C.c_int32),
                ("max_iterations", C.c_int32), ("total_iterations", C.c_int64),
                ("total_dof", C.c_int64), ("device_ms", C.c_double),
                ("worst_true_rel_residual", C.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "pp_device_count": (C.c_int, []),
    "pp_last_error": (C.c_char_p, []),
    "pp_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "pp_ctx_destroy": (C.c_int, [C.c_void_p]),
    "pp_plan_create": (C.c_int, [C.c_void_p, C.POINTER(Mesh), C.POINTER(Material),
                                 C.POINTER(Colouring), C.c_void_p, C.POINTER(C.c_void_p)]),
    "pp_plan_destroy": (C.c_int, [C.c_void_p]),
    "pp_plan_is_structured": (C.c_int, [C.c_void_p]),
    "pp_set_rhs": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pp_solve_resident": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "pp_get_solution": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SolveOpts), C.c_void_p, C.c_void_p,
                           C.POINTER(SolveStats)]),
    "pp_solve_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "pp_apply_k": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "pp_apply_k_colour": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "pp_apply_precond": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_get_block_inverse": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pp_pressure_load": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_int, C.c_void_p]),
    "pp_lognormal_thickness": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                         C.c_double, C.c_uint64, C.c_double, C.c_double,
                                         C.c_void_p, C.c_void_p]),
    "pp_moments": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_max_displacement": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "pp_cnn_create": (C.c_int, [C.c_void_p, C.POINTER(CnnConfig), C.c_uint64,
                                C.POINTER(C.c_void_p)]),
    "pp_cnn_destroy": (C.c_int, [C.c_void_p]),
    "pp_cnn_param_count": (C.c_int64, [C.c_void_p]),
    "pp_cnn_smoother_weight": (C.c_double, [C.c_void_p]),
    "pp_cnn_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pp_cnn_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pp_cnn_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_cnn_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_cnn_mg_levels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "pp_cnn_mg_apply": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "pp_cnn_mg_cycle": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_cnn_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "pp_cnn_gen_error": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "pp_cnn_train": (C.c_int, [C.c_void_p, C.POINTER(TrainOpts), C.POINTER(TrainStats)]),
    "pp_plan_set_learned": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pp_batch_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Mesh), C.POINTER(Material),
                                  C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "pp_batch_destroy": (C.c_int, [C.c_void_p]),
    "pp_batch_plate_count": (C.c_int32, [C.c_void_p]),
    "pp_batch_pressure_load": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p,
                                         C.c_int32, C.c_void_p]),
    "pp_batch_pressure_loads": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pp_batch_set_rhs": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "pp_batch_set_schedule": (C.c_int, [C.c_void_p, C.POINTER(BatchSchedule)]),
    "pp_batch_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(BatchSchedule), C.c_void_p]),
    "pp_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(BatchSchedule),
                                    C.POINTER(BatchStats), C.POINTER(C.c_int32)]),
    "pp_batch_solve": (C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(BatchStats)]),
    "pp_batch_plate_result": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.POINTER(SolveStats)]),
}

_lib = None


class PlatePCGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"platepcg error {code}: {msg}")
        self.code = code


def load(path=None):
    """Load the HIP library once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run __graft_entry__.build() (make -C gpu-gauntlet_amd)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != PP_OK:
        msg = _lib.pp_last_error().decode() if _lib is not None else "?"
        raise PlatePCGError(rc, msg)
    return rc
