"""ctypes binding of libtriblock_hip.so (the C ABI declared in include/triblock_hip.h).

The library is loaded from this package directory only; a missing library is an
error (there is no CPU fallback anywhere in the product path).
"""
import ctypes as C
import os

LIB_NAME = "libtriblock_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

TBS_OK = 0
TBS_ERR_ARG = -1
TBS_ERR_STATE = -2
TBS_ERR_UNSUPPORTED = -3
TBS_ERR_NOMEM = -4
TBS_ERR_IO = -5
TBS_ERR_NUMERIC = -6
TBS_FP64 = 0
TBS_FP32 = 1
TBS_STEP_THOMAS = 0
TBS_STEP_RESIDUAL = 1
TBS_STEP_NORM = 2
TBS_HOST = 0
TBS_DEVICE = 1


class TbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tbs error {code}: {msg}")
        self.code = code


class MeshHost(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nvar", "nelem", "nintface", "eftot", "nline", "nlineelem")] + [
        (n, C.c_void_p) for n in ("epoint", "ef", "fc", "linesize", "linepoint", "lines", "lineface")
    ]


class ServerRequest(C.Structure):
    _fields_ = [("NL", C.c_int), ("NK", C.c_int), ("seed", C.c_uint64), ("shift", C.c_double)]


class PlanInfo(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nvar", "nelem", "nline", "max_faces", "precision")] + [
        ("device_bytes", C.c_size_t),
        ("alg_bytes_per_sweep", C.c_double),
        ("ref_bytes_per_sweep", C.c_double),
        ("thomas_bytes", C.c_double),
        ("residual_bytes", C.c_double),
    ]


class Tuning(C.Structure):
    """tbs_tuning: the tuple the autotuner chose, with its timings."""
    _fields_ = [(n, C.c_int) for n in ("thomas_variant", "thomas_prefetch", "thomas_lds_windows",
                                       "thomas_lines_per_wg", "residual_variant", "residual_waves_per_wg",
                                       "two_stream", "n_thomas", "n_residual")] + [
        ("residual_shapes", C.c_int * 4),
        ("thomas_us", C.c_double * 8),
        ("residual_us", C.c_double * 4),
        ("pass_us", C.c_double * 2),
    ]


P = C.c_void_p
I = C.c_int
L = C.c_long
D = C.c_double
PP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "tbs_last_error": [],
    "tbs_version": [],
    "tbs_supported_nvar": [P, I],
    "tbs_plan_supported_nvar": [P, I],
    "tbs_ctx_create": [I, I, PP],
    "tbs_ctx_destroy": [P],
    "tbs_ctx_set_stream": [P, P],
    "tbs_ctx_stream": [P],
    "tbs_ctx_synchronize": [P],
    "tbs_ctx_nvar": [P],
    "tbs_device_synchronize": [P],
    "tbs_malloc": [P, C.c_size_t, PP],
    "tbs_free": [P, P],
    "tbs_memcpy_h2d": [P, P, P, C.c_size_t],
    "tbs_memcpy_d2h": [P, P, P, C.c_size_t],
    "tbs_memset": [P, P, I, C.c_size_t],
    "tbs_memcpy_d2d": [P, P, P, C.c_size_t],
    "tbs_host_malloc": [P, C.c_size_t, PP],
    "tbs_host_free": [P, P],
    "tbs_event_create": [P, PP],
    "tbs_event_destroy": [P, P],
    "tbs_event_record": [P, P],
    "tbs_event_elapsed_ms": [P, P, C.POINTER(C.c_float)],
    "tbs_stream_ceiling": [P, C.c_size_t, I, P],
    "tbs_copyAtoBjac": [P, I, P, P],
    "tbs_copyAtoB": [P, I, P, P],
    "tbs_addAtoB": [P, I, P, P],
    "tbs_assembleTriBlocks": [P, I, P, P, P, P, P, P, P, P],
    "tbs_lineLU": [P, I, I, I, P, P, P, P, P, P, P, P, P],
    "tbs_triblock_solveDU": [P, I, I, I, I, I, P, P, P, P, P, P, P, P],
    "tbs_triblock_lineRes": [P, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P],
    "tbs_dgemvLU": [P, I, P, P, P],
    "tbs_norm2": [P, L, P, C.POINTER(C.c_double)],
    "tbs_norm2_noalloc": [P, L, P, C.POINTER(C.c_double)],
    "tbs_plan_footprint": [C.POINTER(MeshHost), I, C.POINTER(C.c_size_t)],
    "tbs_plan_create": [P, C.POINTER(MeshHost), I, PP],
    "tbs_plan_destroy": [P],
    "tbs_plan_info_get": [P, C.POINTER(PlanInfo)],
    "tbs_plan_factor": [P, P, P, P],
    "tbs_plans_factor": [P, P, C.c_int, P, P, P],
    "tbs_plan_solve": [P, P, P, I, D, P, C.POINTER(C.c_int), P],
    "tbs_plan_autotune": [P, I, C.POINTER(C.c_int), P, I],
    "tbs_plan_tuning": [P, P],
    "tbs_plan_set_tuning": [P, P],
    "tbs_plan_set_variant": [P, I],
    "tbs_plan_variants": [P, P, I],
    "tbs_plan_begin": [P, P, P],
    "tbs_plan_launch_thomas": [P],
    "tbs_plan_launch_residual": [P],
    "tbs_plan_launch_kernel": [P, I, P, P],
    "tbs_plan_launch_sweeps": [P, I, I],
    "tbs_plan_end": [P, P, P],
    "tbs_plan_history": [P, P, I, C.POINTER(C.c_int)],
    "tbs_plan_internal": [P, I, PP, C.POINTER(C.c_size_t)],
    "tbs_batch_create": [P, P, I, PP],
    "tbs_batch_destroy": [P],
    "tbs_batch_set_plans": [P, P, I],
    "tbs_batch_autotune": [P, I, C.POINTER(C.c_int), P, I],
    "tbs_batch_tuning": [P, P],
    "tbs_batch_set_tuning": [P, P],
    "tbs_batch_solve": [P, P, P, I, D, I, P, P],
    "tbs_batch_launch_sweeps": [P, I, I, D],
    "tbs_batch_states": [P, P],
    "tbs_gmres_create": [P, I, PP],
    "tbs_gmres_destroy": [P],
    "tbs_gmres_solve": [P, P, P, I, D, I, P, I, C.POINTER(C.c_int)],
    "tbs_gmres_stats": [P, P, I],
    "tbs_server_create": [P, C.c_size_t, I, D, I, I, PP],
    "tbs_server_destroy": [P],
    "tbs_server_footprint": [P, C.POINTER(ServerRequest), C.POINTER(C.c_size_t)],
    "tbs_server_submit": [P, C.POINTER(ServerRequest), C.POINTER(C.c_long)],
    "tbs_server_submit_system": [P, C.POINTER(MeshHost), P, P, P, P, I, C.POINTER(C.c_long)],
    "tbs_server_footprint_system": [P, C.POINTER(MeshHost), C.POINTER(C.c_size_t)],
    "tbs_server_timeline": [P, C.c_long, P, I],
    "tbs_server_step": [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "tbs_server_poll": [P, P, I, C.POINTER(C.c_int)],
    "tbs_server_result": [P, C.c_long, P, C.c_size_t, P, I, C.POINTER(C.c_int), C.POINTER(C.c_int), P],
    "tbs_server_release": [P, C.c_long],
    "tbs_server_info": [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_double)],
    "tbs_server_stats": [P, P, I, P, I],
    "tbs_structured_sizes": [I, I, P, P, P, P, P],
    "tbs_structured_mesh": [I, I, P, P, P, P, P, P, P],
    "tbs_ragged_sizes": [I, P, P, P, P],
    "tbs_ragged_mesh": [I, P, P, P, P, P, P, P, P],
    "tbs_batch_shapes": [C.c_uint64, I, L, I, I, I, I, P, P, I, P],
    "tbs_generate_host": [C.c_uint64, I, I, I, D, P, P, P, P],
    "tbs_generate_device": [P, C.c_uint64, I, I, I, D, P, P, P, P],
    "tbs_read_mesh_header": [C.c_char_p, P],
    "tbs_read_mesh": [C.c_char_p, P, P, P, P, P, P, P],
    "tbs_write_mesh": [C.c_char_p, P, P, P, P, P, P, P, P],
    "tbs_mesh_validate": [C.POINTER(MeshHost)],
    "tbs_read_jacobian_header": [C.c_char_p, P],
    "tbs_read_jacobian": [C.c_char_p, P, P, P, P, P],
    "tbs_write_jacobian": [C.c_char_p, I, I, I, P, P, P, P, P],
}
_RESTYPES = {"tbs_last_error": C.c_char_p, "tbs_version": C.c_char_p, "tbs_ctx_stream": C.c_void_p}

_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(rc):
    if rc != TBS_OK:
        raise TbsError(rc, load().tbs_last_error().decode())
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
